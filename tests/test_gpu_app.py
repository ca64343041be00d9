"""A caller shaped like the reference's command-line drivers
(app/implicit_arrangement.cpp, app/material_interface.cpp), compiled against
this build's headers only (tests/native/app_caller.cpp): configuration file,
lookup-table switch, generate_tet_mesh or load_tet_mesh, load_functions on the
GPU, the drop-in entry point, and mesh.json / stats.json.  The files it
writes equal, byte for byte, the files the oracle's hot path and the same host
stages and writers produce (tests/native/ka_glue.cpp)."""
import json
import os
import subprocess

import numpy as np
import pytest

import support as S
from test_io_formats import ka_write, write_mesh_json

pytestmark = pytest.mark.gpu

APP = os.path.join(S.ROOT, "tests", "native", "app_caller")
KA = json.load(open(os.path.join(S.GOLDEN, "known_answers.json")))


def _run(mode, cfg, tmp_path, robust=False):
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    cfg = dict(cfg, outputDir="out", useSecondaryLookup=True, useTopoRayShooting=True)
    cfg.setdefault("useLookup", True)
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    print("app_caller", mode, json.dumps(cfg), flush=True)
    r = subprocess.run([APP, mode, str(tmp_path / "config.json")] + (["-R"] if robust else []), capture_output=True,
                       text=True, timeout=300)
    return r, out


def _reference_files(mi, pts, tets, funcs, tmp_path):
    ref = tmp_path / "ref"
    ref.mkdir()
    types, params = S.load_ref_funcs(funcs)
    assert ka_write(mi, pts, tets, S.evaluate(types, params, pts), ref) >= 0
    return ref


@pytest.mark.parametrize("case", ["3-sphere-4", "2-planesphere"])
@pytest.mark.parametrize("lookup", [True, False])
def test_app_implicit_arrangement_on_a_generated_grid(case, lookup, tmp_path):
    ka = next(c for c in KA["ia"] if c["funcs"] == case + ".json" and not c.get("negate"))
    cfg = {"gridResolution": 101, "gridBbox": [[-1, -1, -1], [1, 1, 1]],
           "funcFile": os.path.join(S.REF_TESTS, ka["funcs"]), "useLookup": lookup}
    r, out = _run("ia", cfg, tmp_path)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert json.loads(r.stdout.strip().splitlines()[-1])["ok"]
    mesh = json.load(open(out / "mesh.json"))
    assert len(mesh["patches"]) == ka["patches"] and len(mesh["cells"]) == ka["cells"]
    assert mesh["patches_label"] == [ka["patch_label"]]
    pts, tets = S.grid(101)
    ref = _reference_files(0, pts, tets, ka["funcs"], tmp_path)
    assert (out / "mesh.json").read_bytes() == (ref / "mesh.json").read_bytes()
    st = json.load(open(out / "stats.json"))
    rst = json.load(open(ref / "stats.json"))
    for k in ("num_patches", "num_chains", "num_shells", "num_cells", "num_iso_edges"):
        assert st[k] == rst[k], k


def test_app_material_interface_on_a_mesh_file(tmp_path):
    ka = next(c for c in KA["mi"] if c.get("mesh"))
    m = np.load(os.path.join(S.REF_TESTS, ka["mesh"]))
    write_mesh_json(tmp_path / "mesh_in.json", m["pts"], m["tets"])
    cfg = {"tetMeshFile": "mesh_in.json", "funcFile": os.path.join(S.REF_TESTS, ka["funcs"])}
    r, out = _run("mi", cfg, tmp_path)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    mesh = json.load(open(out / "mesh.json"))
    assert len(mesh["cells"]) == ka["cells"] and mesh["patches_label"] == [ka["patch_label"]]
    ref = _reference_files(1, m["pts"], m["tets"].astype(np.uint64), ka["funcs"], tmp_path)
    assert (out / "mesh.json").read_bytes() == (ref / "mesh.json").read_bytes()


@pytest.mark.parametrize("mode", ["ia", "mi"])
def test_app_robust_test(mode, tmp_path):
    funcs = "3-sphere-1.json" if mode == "ia" else "8-sphere.json"
    cfg = {"gridResolution": 40, "gridBbox": [[-1, -1, -1], [1, 1, 1]], "funcFile": os.path.join(S.REF_TESTS, funcs)}
    r, out = _run(mode, cfg, tmp_path, robust=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "success." in r.stdout.splitlines()
    assert not (out / "mesh.json").exists()
    st = json.load(open(out / "stats.json"))
    pts, tets = S.grid(40)
    types, params = S.load_ref_funcs(funcs)
    ref = S.OracleRun(int(mode == "mi"), pts, tets, S.evaluate(types, params, pts))
    assert st["num_intersecting_tet"] == ref.num_intersecting
    keys = ["num_1_func", "num_2_func", "num_more_func"] if mode == "ia" else ["num_2_func", "num_3_func",
                                                                              "num_more_func"]
    assert [st[k] for k in keys] == ref.buckets


def test_app_failing_reference_case(tmp_path):
    """tests/test_implicit_networks.cpp:734-812: material_interface fails in the
    graph stage on mesh_fail.json; the driver returns -1 with the reference's
    message."""
    m = np.load(os.path.join(S.REF_TESTS, "mesh_fail.npz"))
    write_mesh_json(tmp_path / "mesh_in.json", m["pts"], m["tets"])
    cfg = {"tetMeshFile": "mesh_in.json", "funcFile": os.path.join(S.REF_TESTS, "8-sphere.json")}
    r, out = _run("mi", cfg, tmp_path)
    assert r.returncode == 256 - 1
    assert ("order patches failed: More than one tets containing the iso-faces incident to the iso-edge."
            in r.stdout.splitlines())
