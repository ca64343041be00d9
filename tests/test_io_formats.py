"""The reference's file formats and driver pieces, on the CPU:
generate_tet_mesh and load_tet_mesh (src/io.cpp:65-152), parse_config_file
(src/io.cpp:19-63), the mesh.json / timings.json / stats.json writers
(src/io.cpp:154-312, :556-590) and the -R verdict
(src/implicit_arrangement.cpp:329-343).  The writers are checked byte for
byte against nlohmann::json (the reference's JSON library; the image carries
its 3.1.1 header, the reference pins 3.7.3) where the header is present, and
by reading the files back everywhere."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import support as S
import workloads as W

NATIVE = os.path.join(S.ROOT, "tests", "native")
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _ka():
    lib = ctypes.CDLL(os.path.join(NATIVE, "libka.so"))
    lib.ka_write.restype = ctypes.c_int
    lib.ka_write.argtypes = [ctypes.c_int, _f64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f64p, ctypes.c_uint64,
                             ctypes.c_char_p]
    lib.io_generate.argtypes = [ctypes.c_uint64, _f64p, _f64p, _f64p, _u64p]
    lib.io_load.argtypes = [ctypes.c_char_p, _u64p, _u64p]
    lib.io_mesh_copy.argtypes = [_f64p, _u64p]
    lib.io_config.restype = ctypes.c_char_p
    lib.io_config.argtypes = [ctypes.c_char_p]
    return lib


def ka_write(mi, pts, tets, vals, out_dir):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    return _ka().ka_write(int(mi), pts.ctypes.data_as(_f64p), pts.shape[0], tets.ctypes.data_as(_u64p), tets.shape[0],
                          vals.ctypes.data_as(_f64p), vals.shape[1], str(out_dir).encode())


def write_mesh_json(path, pts, tets):
    with open(path, "w") as f:
        json.dump([pts.tolist(), [[int(x) for x in t] for t in tets]], f)


@pytest.mark.parametrize("res,lo,hi", [(3, (-1, -1, -1), (1, 1, 1)), (7, (0.0, -2.0, 0.5), (1.0, 2.0, 0.75))])
def test_generate_tet_mesh_matches_the_grid(res, lo, hi):
    N = res + 1
    pts = np.zeros((N ** 3, 3))
    tets = np.zeros((5 * res ** 3, 4), dtype=np.uint64)
    lo_a, hi_a = np.asarray(lo, float), np.asarray(hi, float)
    assert _ka().io_generate(res, lo_a.ctypes.data_as(_f64p), hi_a.ctypes.data_as(_f64p), pts.ctypes.data_as(_f64p),
                             tets.ctypes.data_as(_u64p)) == 0
    p2, t2 = S.grid(res, lo, hi)
    np.testing.assert_array_equal(pts.view(np.uint64), p2.view(np.uint64))
    np.testing.assert_array_equal(tets, t2)


def test_load_tet_mesh_reads_the_reference_mesh(tmp_path):
    m = np.load(os.path.join(S.REF_TESTS, "mesh.npz"))
    path = tmp_path / "mesh.json"
    write_mesh_json(path, m["pts"], m["tets"])
    V, T = ctypes.c_uint64(), ctypes.c_uint64()
    lib = _ka()
    assert lib.io_load(str(path).encode(), ctypes.byref(V), ctypes.byref(T)) == 0
    pts = np.zeros((V.value, 3))
    tets = np.zeros((T.value, 4), dtype=np.uint64)
    lib.io_mesh_copy(pts.ctypes.data_as(_f64p), tets.ctypes.data_as(_u64p))
    np.testing.assert_array_equal(pts.view(np.uint64), m["pts"].view(np.uint64))
    np.testing.assert_array_equal(tets, m["tets"].astype(np.uint64))
    assert lib.io_load(str(tmp_path / "missing.json").encode(), ctypes.byref(V), ctypes.byref(T)) == -1


def test_parse_config_file_resolves_relative_paths(tmp_path):
    cfg = {"gridResolution": 32, "gridBbox": [[-1, -0.5, 0], [1, 0.5, 2]], "funcFile": "funcs/f.json",
           "outputDir": "../out", "useLookup": True, "useSecondaryLookup": False, "useTopoRayShooting": True}
    (tmp_path / "cfg").mkdir()
    p = tmp_path / "cfg" / "config.json"
    p.write_text(json.dumps(cfg))
    got = json.loads(_ka().io_config(str(p).encode()).decode())
    assert got["tet_mesh_file"] == "" and got["res"] == 32
    assert got["bbox"] == [-1, -0.5, 0, 1, 0.5, 2]
    assert got["func_file"] == str(tmp_path / "cfg" / "funcs" / "f.json")
    assert os.path.normpath(got["output_dir"]) == str(tmp_path / "out")
    assert (got["use_lookup"], got["use_secondary_lookup"], got["use_topo_ray_shooting"]) == (1, 0, 1)
    cfg2 = dict(cfg, tetMeshFile="m.json")
    p.write_text(json.dumps(cfg2))
    got = json.loads(_ka().io_config(str(p).encode()).decode())
    assert got["tet_mesh_file"] == str(tmp_path / "cfg" / "m.json")
    got = json.loads(_ka().io_config(str(tmp_path / "none.json").encode()).decode())
    assert got["error"] == "Config file does not exist!"


def _case(which):
    if which == "ia":
        res, fs = W.config1()
        mi = 0
    elif which == "ia_degenerate":
        res, fs = W.config4()
        res, mi = 12, 0
    else:
        _, fs = W.config3()
        res, mi = 16, 1
    types, params = fs.arrays()
    pts, tets = S.grid(res)
    return mi, pts, tets, S.evaluate(types, params, pts)


@pytest.mark.parametrize("which", ["ia", "ia_degenerate", "mi"])
def test_mesh_json_holds_the_result(which, tmp_path):
    mi, pts, tets, vals = _case(which)
    assert ka_write(mi, pts, tets, vals, tmp_path) >= 0
    d = json.load(open(tmp_path / "mesh.json"))
    assert sorted(d) == ["cells", "cells_label", "chains", "corners", "edges", "faces", "patches", "patches_label",
                         "points", "shells"]
    ref = S.OracleRun(mi, pts, tets, vals)
    np.testing.assert_array_equal(np.asarray(d["points"]).view(np.uint64), ref.xyz.view(np.uint64))
    faces = [ref.face_verts[ref.face_start[i]:ref.face_start[i + 1]].tolist() for i in range(len(ref.face_func))]
    assert d["faces"] == faces
    out = S.known_answer_run(mi, pts, tets, vals)
    assert len(d["patches"] or []) == out["patches"] and len(d["cells"] or []) == out["cells"]
    assert d["patches_label"] == [out["patch_label"]]
    st = json.load(open(tmp_path / "stats.json"))
    assert st["num_pts"] == len(pts)
    assert list(json.load(open(tmp_path / "timings.json"))) == sorted(json.load(open(tmp_path / "timings.json")))


@pytest.mark.parametrize("which", ["ia", "ia_degenerate", "mi"])
def test_writers_match_nlohmann_byte_for_byte(which, tmp_path):
    mi, pts, tets, vals = _case(which)
    rc = ka_write(mi, pts, tets, vals, tmp_path)
    if rc == 0:
        pytest.skip("nlohmann::json is not in this image; the checker was not built")
    assert rc == 1
    for name in ("mesh", "timings", "stats"):
        a = (tmp_path / (name + ".json")).read_bytes()
        b = (tmp_path / (name + "_nlohmann.json")).read_bytes()
        assert a == b, name


def test_robust_verdict_messages(capfd):
    lib = ctypes.CDLL(os.path.join(NATIVE, "libdropin.so"))
    lib.verdict_probe.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, _u64p, ctypes.c_uint64]
    a = np.array([4, 5, 2, 7, 9, 3], dtype=np.uint64)
    b = a.copy()
    b[4] = 8  # one tet's face count differs
    p = lambda x: x.ctypes.data_as(_u64p)  # noqa: E731
    cases = [((1, 1, a, a), 1, "success."), ((1, 1, a, b), 0, "type 1 failure (inconsistency)."),
             ((0, 1, a, a), 0, "type 2 failure (crash in the normal order)."),
             ((1, 0, a, a), 0, "type 3 failure (crash in the reverse order)."),
             ((0, 0, a, b), 0, "type 2 failure (crash in the normal order).")]
    for (f, r, x, y), want, msg in cases:
        assert lib.verdict_probe(f, r, p(x), p(y), len(a)) == want
        assert capfd.readouterr().out.strip() == msg


def test_app_rejects_an_unknown_function_type(tmp_path):
    """The shaped caller stops with the reference's exit code -2 when
    load_functions fails (app/implicit_arrangement.cpp:59-64)."""
    (tmp_path / "f.json").write_text(json.dumps([{"type": "sphere", "center": [0, 0, 0], "radius": 0.5},
                                                 {"type": "circle", "center": [0, 0, 0], "axis_vector": [0, 0, 1],
                                                  "radius": 0.3}]))
    cfg = {"gridResolution": 4, "gridBbox": [[-1, -1, -1], [1, 1, 1]], "funcFile": "f.json", "outputDir": ".",
           "useLookup": True, "useSecondaryLookup": True, "useTopoRayShooting": True}
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    r = subprocess.run([os.path.join(NATIVE, "app_caller"), "ia", str(tmp_path / "c.json")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 256 - 2
    assert "unsupported function type \"circle\"" in r.stdout and "function loading failed." in r.stdout
