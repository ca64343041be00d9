"""Device pipeline (through the C ABI) against the oracle on the same inputs:
grid points and function values bit for bit, the active-function CSR, the
bucket counts, iso-vertices (records and coordinates) and iso-faces (vertex
cycles, tet-face occurrences, function labels) bit for bit."""
import numpy as np
import pytest

import support as S
import workloads as W

pytestmark = pytest.mark.gpu


def _compare(dev, ref):
    s = dev.sizes
    assert dev.num_degenerate == ref.num_degenerate
    assert dev.num_intersecting == ref.num_intersecting
    assert dev.buckets == ref.buckets
    fit, st = dev.csr()
    np.testing.assert_array_equal(st, ref.start_index)
    np.testing.assert_array_equal(fit, ref.func_in_tet)
    rec, xyz = dev.verts()
    assert rec.shape == ref.verts.shape, (rec.shape, ref.verts.shape)
    np.testing.assert_array_equal(rec, ref.verts)
    # bit-exact coordinates (same expression order, IEEE division)
    np.testing.assert_array_equal(xyz.view(np.uint64), ref.xyz.view(np.uint64))
    vs, v, ts, tf, fn = dev.faces()
    np.testing.assert_array_equal(vs, ref.face_start)
    np.testing.assert_array_equal(v, ref.face_verts)
    np.testing.assert_array_equal(ts, ref.tf_start)
    np.testing.assert_array_equal(tf, ref.tf)
    np.testing.assert_array_equal(fn, ref.face_func)


def test_grid_points(ctx):
    for res in (20, 32):
        g = ctx.grid_regular(res)
        pts, _ = S.grid(res)
        np.testing.assert_array_equal(g.points().view(np.uint64), pts.view(np.uint64))
        g.release()


def test_eval_bit_exact(ctx):
    res, fs = W.config2()
    res = 24
    g = ctx.grid_regular(res)
    types, params = fs.arrays()
    # squared and plain spheres both
    fs2 = W.mixed(np.random.default_rng(9), 2, 3, 2, 2)
    fs2.sphere([0.1, 0.2, -0.3], 0.45, squared=True)
    for t, p in (fs.arrays(), fs2.arrays()):
        vals = g.eval(t, p)
        pts, _ = S.grid(res)
        ref = S.evaluate(t, p, pts)
        np.testing.assert_array_equal(vals.download().view(np.uint64), ref.view(np.uint64))
        vals.release()
    g.release()


@pytest.mark.parametrize("which", ["config1", "config2_small", "config4", "config4_shuffled"])
def test_pipeline_ia_matches_oracle(ctx, which):
    if which == "config1":
        res, fs = W.config1()
    elif which == "config2_small":
        _, fs = W.config2()
        res = 40
    elif which == "config4":
        res, fs = W.config4()
    else:
        res, fs = W.config4(shuffle_seed=44)
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(0)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    _compare(dev, ref)
    dev.release()
    vals.release()
    g.release()


def test_pipeline_ia_explicit_mesh(ctx):
    m = np.load(S.REF_TESTS + "/mesh.npz")
    pts, tets = m["pts"], m["tets"].astype(np.uint64)
    types, params = S.load_ref_funcs("8-sphere.json")
    g = ctx.grid_explicit(pts, tets)
    vals = g.upload(S.evaluate(types, params, pts))
    dev = vals.run(0)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    _compare(dev, ref)


@pytest.mark.parametrize("which", ["config3_small", "config3", "config4", "config4_shuffled", "config1"])
def test_pipeline_mi_matches_oracle(ctx, which):
    """Material interface: highest material, filter, MI cut, extract_MI_mesh
    (lazy corner vertices, tet-boundary face pairing) and coordinates."""
    import rin

    if which == "config3_small":
        _, fs = W.config3()
        res = 24
    elif which == "config3":
        res, fs = W.config3()
    elif which == "config4":
        res, fs = W.config4()
    elif which == "config4_shuffled":
        res, fs = W.config4(shuffle_seed=44)
    else:
        res, fs = W.config1()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(rin.FLAG_MI)
    pts, tets = S.grid(res)
    ref = S.OracleRun(1, pts, tets, S.evaluate(types, params, pts))
    assert dev.sizes.n_faces > 0
    _compare(dev, ref)
    dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("which", ["config3_small", "config4_shuffled"])
def test_mi_two_material_table(ctx, which):
    """Two-material tets take their cut record from the sign-pattern table:
    every record equals the oracle's cut of the same tet, and the outputs equal
    those of the general cut for every tet (lookups off)."""
    import rin

    if which == "config3_small":
        _, fs = W.config3()
        res = 24
    else:
        res, fs = W.config4(shuffle_seed=44)
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    on = vals.run(rin.FLAG_MI | rin.FLAG_KEEP_RECORDS)
    off = vals.run(rin.FLAG_MI | rin.FLAG_NO_LOOKUP | rin.FLAG_KEEP_RECORDS)
    # the batched form keeps no copy: a table tet's record is the table's
    batched = vals.run(rin.FLAG_MI | rin.FLAG_BATCHED | rin.FLAG_KEEP_RECORDS)
    pts, tets = S.grid(res)
    ref = S.OracleRun(1, pts, tets, S.evaluate(types, params, pts))
    assert on.buckets[0] > 0  # two-material tets present
    cap = len(ref.words) + 16
    st, w = on.records(ref.active, cap)
    np.testing.assert_array_equal(st, ref.rec_start)
    np.testing.assert_array_equal(w, ref.words)
    st2, w2 = off.records(ref.active, cap)
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(w2, w)
    st3, w3 = batched.records(ref.active, cap)
    np.testing.assert_array_equal(st3, st)
    np.testing.assert_array_equal(w3, w)
    _compare(on, ref)
    _compare(off, ref)
    _compare(batched, ref)
    for r in (on, off, batched):
        r.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("which", ["config1", "config4_shuffled"])
def test_plain_ia_tables(ctx, which):
    """The plain form's one- and two-function tets take their records from
    the k = 1 and k = 2 tables: every record equals the oracle's cut of the
    same tet and the general cut's (lookups off), and both runs give the
    oracle's outputs."""
    import rin

    if which == "config1":
        res, fs = W.config1()
    else:
        res, fs = W.config4(shuffle_seed=44)
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    on = vals.run(rin.FLAG_KEEP_RECORDS)
    off = vals.run(rin.FLAG_NO_LOOKUP | rin.FLAG_KEEP_RECORDS)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    assert on.buckets[0] > 0 and on.buckets[1] > 0  # one- and two-function tets present
    cap = len(ref.words) + 16
    st, w = on.records(ref.active, cap)
    np.testing.assert_array_equal(st, ref.rec_start)
    np.testing.assert_array_equal(w, ref.words)
    st2, w2 = off.records(ref.active, cap)
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(w2, w)
    _compare(on, ref)
    _compare(off, ref)
    for r in (on, off):
        r.release()
    vals.release()
    g.release()


def test_pipeline_mi_explicit_mesh(ctx):
    """The reference's unstructured eight-sphere material-interface case."""
    import rin

    m = np.load(S.REF_TESTS + "/mesh.npz")
    pts, tets = m["pts"], m["tets"].astype(np.uint64)
    types, params = S.load_ref_funcs("8-sphere.json")
    g = ctx.grid_explicit(pts, tets)
    vals = g.upload(S.evaluate(types, params, pts))
    dev = vals.run(rin.FLAG_MI)
    ref = S.OracleRun(1, pts, tets, S.evaluate(types, params, pts))
    _compare(dev, ref)


@pytest.mark.parametrize("form", ["auto", "wave", "wave16", "thread", "table", "wave-scratch", "wave16-scratch"])
@pytest.mark.parametrize("which", ["config1", "config2_small", "config4"])
def test_batched_matches_oracle_and_replays(ctx, which, form, monkeypatch):
    """The batched form (k=1 table, bucket kernels, captured graph) gives the
    oracle's output bit for bit, on the first (capturing) run and on graph
    replays with plain-form runs between them; the bucket cut kernels are the
    autotuned choice, or forced to the wave-per-tet or thread-per-tet form, or
    k = 2 by the sign-vector table.  The lane-group cuts of 3..8 functions
    start from the k = 2 table's complex after two planes, or ("-scratch")
    insert every plane themselves."""
    import rin

    if form.endswith("-scratch"):
        monkeypatch.setenv("RIN_K2_START", "0")
        form = form[: -len("-scratch")]
    if form != "auto":
        monkeypatch.setenv("RIN_CUT_FORM", form)

    if which == "config1":
        res, fs = W.config1()
    elif which == "config2_small":
        _, fs = W.config2()
        res = 48
    else:
        res, fs = W.config4()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    for rep in range(3):
        dev = vals.run(rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
        plain = vals.run(0)
        plain.release()
    # a held result forces a second workspace; both stay correct
    a = vals.run(rin.FLAG_BATCHED)
    b = vals.run(rin.FLAG_BATCHED)
    _compare(a, ref)
    _compare(b, ref)
    a.release()
    b.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("seed", [101, 102, 103, 104, 105, 106])
def test_random_sets_all_forms(ctx, seed):
    """Seeded random function sets (2..14 mixed functions, resolutions 9..27,
    some with exactly repeated functions): the batched and plain implicit
    arrangements and the material interface match the oracle bit for bit."""
    import rin

    rng = np.random.default_rng(seed)
    n = [int(x) for x in rng.integers(0, 5, 4)]
    n[0] += 1
    fs = W.mixed(rng, *n)
    if seed % 2:
        # a repeated function: coincident planes and groups in the cuts
        t0, p0 = fs.arrays()
        fs.types.append(int(t0[0]))
        fs.params.append(list(p0[0]))
    res = int(rng.integers(9, 28))
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    hv = S.evaluate(types, params, pts)
    ref = S.OracleRun(0, pts, tets, hv)
    for flags in (rin.FLAG_BATCHED, 0):
        dev = vals.run(flags)
        _compare(dev, ref)
        dev.release()
    if len(types) <= 12:
        ref_mi = S.OracleRun(1, pts, tets, hv)
        dev = vals.run(rin.FLAG_MI)
        _compare(dev, ref_mi)
        dev.release()
    vals.release()
    g.release()


def test_batched_explicit_mesh(ctx):
    """The batched form on an unstructured mesh (no star tables: every
    occurrence takes the hash path, ids by a global scan), across replays."""
    import rin

    m = np.load(S.REF_TESTS + "/mesh.npz")
    pts, tets = m["pts"], m["tets"].astype(np.uint64)
    types, params = S.load_ref_funcs("8-sphere.json")
    g = ctx.grid_explicit(pts, tets)
    vals = g.upload(S.evaluate(types, params, pts))
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    for rep in range(3):
        dev = vals.run(rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("mi", [0, 1])
@pytest.mark.parametrize("batched", [False, True])
def test_large_complexes_are_cut_again(ctx, mi, batched):
    """Many functions crossing inside a few large tets: complexes outgrow the
    bucket kernels' capacities, are cut again with the large-capacity kernel
    (the batched form hands the run to the plain form) and still match the
    oracle bit for bit."""
    import rin

    rng = np.random.default_rng(31 + mi)
    fs = W.FuncSet()
    for _ in range(8):
        if mi:
            fs.sphere(rng.uniform(-0.3, 0.3, 3), rng.uniform(0.0, 0.1))
        else:
            fs.plane(rng.uniform(-0.2, 0.2, 3), rng.standard_normal(3))
    types, params = fs.arrays()
    res = 2
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    flags = (rin.FLAG_MI if mi else 0) | (rin.FLAG_BATCHED if batched and not mi else 0)
    dev = vals.run(flags)
    pts, tets = S.grid(res)
    ref = S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))
    _compare(dev, ref)
    dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("tile", ["256", "512", "1024"])
@pytest.mark.parametrize("which", ["config1", "config4"])
def test_batched_cube_tiles(ctx, which, tile, monkeypatch):
    """Every cube tile the autotuner can pick for the count and fill passes
    (256, 512 or 1024 cubes a workgroup) gives the oracle's output bit for bit,
    on the capturing run and on a replay."""
    import rin

    monkeypatch.setenv("RIN_CUBE_TILE", tile)
    res, fs = W.config1() if which == "config1" else W.config4()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    for rep in range(2):
        dev = vals.run(rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("form", ["wave16", "wave"])
def test_batched_bucket_lists_outgrow_the_plan(ctx, form, monkeypatch):
    """A second value set of the same size with many more tets of 2..5
    active functions than the first overflows the plan's bucket capacities
    (the count pass's k = 3..4 list, with its k = 4 front and k = 3 back
    parts, included): the run grows the workspace and still matches the
    oracle; both sets, in turn, keep matching it afterwards.  Counts per k,
    res 32: first set 1804 / 47 / 0 / 0 tets of k = 2 / 3 / 4 / 5, second
    set 4372 / 446 / 44 / 12."""
    import rin

    monkeypatch.setenv("RIN_CUT_FORM", form)
    res = 32
    few = W.FuncSet()
    few.sphere([0.1, -0.2, 0.15], 0.5)
    few.sphere([-0.2, 0.1, 0.0], 0.45)
    for i in range(6):
        few.plane([-0.75 + 0.3 * i, 0, 0], [1, 0, 0])
    many = W.mixed(np.random.default_rng(33), 2, 2, 2, 2)
    g = ctx.grid_regular(res)
    pts, tets = S.grid(res)
    sets = []
    for fs in (few, many):
        types, params = fs.arrays()
        sets.append((g.eval(types, params), S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))))
    assert sets[0][1].buckets != sets[1][1].buckets
    for i in (0, 1, 0, 1):
        vals, ref = sets[i]
        dev = vals.run(rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
    for vals, _ in sets:
        vals.release()
    g.release()


@pytest.mark.parametrize("form", ["wave16", "wave", "thread", "plain"])
def test_sixty_four_functions(ctx, form, monkeypatch):
    """64 functions, the widest sign masks (every bit a function): the
    batched form's count-pass lists, the k = 2 lookups and the cuts, which
    take a tet's functions from its corner masks, and the plain form match
    the oracle.  64 small spheres at res 24: 14898 / 4287 / 1160 / 380 / 128
    / 23 / 8 tets of k = 1..7, every function active somewhere."""
    import rin

    if form != "plain":
        monkeypatch.setenv("RIN_CUT_FORM", form)
    res = 24
    rng = np.random.default_rng(64)
    fs = W.FuncSet()
    for _ in range(64):
        fs.sphere(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.1, 0.3))
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    assert len(np.unique(ref.func_in_tet)) == 64
    flags = 0 if form == "plain" else rin.FLAG_BATCHED
    for _ in range(2):
        dev = vals.run(flags)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("batched", [True, False])
def test_mi_sixty_four_materials(ctx, batched):
    """64 materials (every mask bit a material) through the material
    interface, batched and plain, against the oracle: Voronoi-like
    materials -|x - p_i| + w_i at res 24, 16626 / 6337 / 3089 tets of 2 / 3
    / 4 or more materials (up to 9: the batched form takes those with its
    large-capacity cut and stays one graph)."""
    import rin

    res = 24
    rng = np.random.default_rng(164)
    fs = W.FuncSet()
    for _ in range(64):
        fs.sphere(rng.uniform(-0.9, 0.9, 3), rng.uniform(0.0, 0.1))
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(1, pts, tets, S.evaluate(types, params, pts))
    assert ref.num_intersecting > 0
    flags = rin.FLAG_MI | (rin.FLAG_BATCHED if batched else 0)
    for _ in range(2):
        dev = vals.run(flags)
        if batched:
            label = dev.stage_times()[0][0]
            assert label == "batched path (one graph)", label
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("mi", [0, 1])
def test_no_active_tets_then_some(ctx, mi):
    """Functions that cross no tet (planes outside the grid's box; for the
    material interface, one material above the others everywhere) give the
    oracle's empty output in every form, on the capturing run and a replay;
    a value set of the same size with active tets then grows the same plan
    from empty buckets and still matches."""
    import rin

    res = 16
    empty = W.FuncSet()
    for j in range(4):
        empty.plane([1.5 + 0.1 * j, 0, 0], [1, 0.2 * j, 0])
    if mi:
        # material 0 is above the others everywhere in the box
        empty = W.FuncSet()
        empty.plane([5.0, 0, 0], [-1, 0, 0])
        for j in range(3):
            empty.plane([-3.0 - j, 0, 0], [-1, 0, 0])
    some = W.mixed(np.random.default_rng(17), 1, 1, 1, 1)
    g = ctx.grid_regular(res)
    pts, tets = S.grid(res)
    sets = []
    for fs in (empty, some):
        types, params = fs.arrays()
        sets.append((g.eval(types, params), S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))))
    assert sets[0][1].num_intersecting == 0 and sets[1][1].num_intersecting > 0
    base = rin.FLAG_MI if mi else 0
    for flags in (base, base | rin.FLAG_BATCHED):
        for i in (0, 0, 1, 0):
            vals, ref = sets[i]
            dev = vals.run(flags)
            _compare(dev, ref)
            dev.release()
    for vals, _ in sets:
        vals.release()
    g.release()


@pytest.mark.parametrize("res", [1, 2, 3])
@pytest.mark.parametrize("mi", [0, 1])
def test_tiny_grids(ctx, res, mi):
    """Grids of one, eight and 27 cubes (a single, partly filled cube tile)
    with functions crossing them, in every form, against the oracle."""
    import rin

    fs = W.mixed(np.random.default_rng(100 + res), 2, 2, 1, 1)
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))
    assert ref.num_intersecting > 0
    base = rin.FLAG_MI if mi else 0
    for flags in (base, base | rin.FLAG_BATCHED, base | rin.FLAG_BATCHED):
        dev = vals.run(flags)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("form", ["batched", "plain", "batched-scratch"])
def test_headline_workload_full_size(ctx, form, monkeypatch):
    """The benchmark workload itself: 101^3 vertices (5,000,000 tets), 20
    functions (6 planes, 6 spheres, 4 cylinders, 4 tori), in both forms, bit
    for bit against the oracle (counts, CSR, vertices and coordinates, faces);
    "-scratch": the lane-group cuts insert every plane themselves."""
    import rin

    if form.endswith("-scratch"):
        monkeypatch.setenv("RIN_K2_START", "0")
        form = form[: -len("-scratch")]

    res, fs = W.config2()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))
    flags = rin.FLAG_BATCHED if form == "batched" else 0
    for rep in range(2 if form == "batched" else 1):  # the capturing run, then a replay
        dev = vals.run(flags)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


@pytest.mark.parametrize("mi", [0, 1])
def test_failing_reference_mesh_hot_path(ctx, mi):
    """tests/test_implicit_networks.cpp:734-812 (mesh_fail.json, 8-sphere.json):
    the reference's graph stage fails there, the hot path's outputs must still
    equal the oracle's."""
    import rin

    m = np.load(S.REF_TESTS + "/mesh_fail.npz")
    pts, tets = m["pts"], m["tets"].astype(np.uint64)
    types, params = S.load_ref_funcs("8-sphere.json")
    hv = S.evaluate(types, params, pts)
    g = ctx.grid_explicit(pts, tets)
    vals = g.upload(hv)
    ref = S.OracleRun(mi, pts, tets, hv)
    for flags in ((rin.FLAG_MI,) if mi else (0, rin.FLAG_BATCHED)):
        dev = vals.run(flags)
        _compare(dev, ref)
        dev.release()
    vals.release()
    g.release()


def _record_counts(ref):
    """(vertices, faces, cells) of every active tet from the oracle's records."""
    out = np.zeros((len(ref.active), 3), dtype=np.uint64)
    for a in range(len(ref.active)):
        w = ref.words[ref.rec_start[a]:ref.rec_start[a + 1]]
        out[a] = (w[2], w[3], w[4])
    return out


@pytest.mark.parametrize("mi,which", [(0, "config1"), (0, "config4"), (1, "config3_small"), (1, "config4")])
@pytest.mark.parametrize("lookup", [True, False])
def test_robust_check_counts_match_oracle(ctx, mi, which, lookup):
    """-R (src/implicit_arrangement.cpp:136-243, src/material_interface.cpp:174-287):
    the per-tet (vertex, face, cell) counts of the forward and the reversed
    insertion order equal the oracle's cuts in the same orders, and the run
    stops after the cut."""
    import rin

    if which == "config1":
        res, fs = W.config1()
    elif which == "config4":
        res, fs = W.config4(shuffle_seed=44)
        res = 16
    else:
        _, fs = W.config3()
        res = 24
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    hv = S.evaluate(types, params, pts)
    base = (rin.FLAG_MI if mi else 0) | (0 if lookup else rin.FLAG_NO_LOOKUP) | rin.FLAG_CUT_ONLY
    for rev in (False, True):
        ref = S.OracleRun(mi, pts, tets, hv, flags=2 if rev else 0)
        dev = vals.run(base | (rin.FLAG_REVERSE if rev else 0))
        assert dev.num_intersecting == ref.num_intersecting and dev.buckets == ref.buckets
        assert dev.num_degenerate == ref.num_degenerate
        assert dev.sizes.n_verts == 0 and dev.sizes.n_faces == 0
        np.testing.assert_array_equal(dev.cut_counts(), _record_counts(ref))
        dev.release()
    vals.release()
    g.release()


def test_values_from_another_grid_are_rejected(ctx):
    """Values carry the grid they were made on; a run on another grid of the
    same vertex count is an argument error, never a replay of the first
    grid's graph."""
    import rin

    m = np.load(S.REF_TESTS + "/mesh.npz")
    pts, tets = m["pts"], m["tets"].astype(np.uint64)
    types, params = S.load_ref_funcs("8-sphere.json")
    hv = S.evaluate(types, params, pts)
    g1 = ctx.grid_explicit(pts, tets)
    g2 = ctx.grid_explicit(pts, tets[::-1].copy())  # same V, other tets
    v1 = g1.upload(hv)
    v1.run(rin.FLAG_BATCHED).release()  # captures g1's graph
    with pytest.raises(rin.RinError) as e:
        v1.grid = g2
        v1.run(rin.FLAG_BATCHED)
    assert e.value.code == -1
    v1.grid = g1
    v2 = g2.upload(hv)
    ref2 = S.OracleRun(0, pts, tets[::-1].copy(), hv)
    dev = v2.run(rin.FLAG_BATCHED)
    _compare(dev, ref2)
    dev.release()
    for o in (v1, v2, g1, g2):
        o.release()


def test_value_sets_share_the_grid_plan(ctx):
    """Several value sets of the same size on one grid share the batched
    form's workspace and graph: switching sets re-records the graph and
    updates it in place; every run matches the oracle for its own values."""
    import rin

    res = 40
    g = ctx.grid_regular(res)
    pts, tets = S.grid(res)
    sets = []
    for seed in (21, 22, 23):
        fs = W.mixed(np.random.default_rng(seed), 4, 4, 2, 2)
        types, params = fs.arrays()
        vals = g.eval(types, params)
        sets.append((vals, S.OracleRun(0, pts, tets, S.evaluate(types, params, pts))))
    for i in (0, 1, 2, 1, 0, 2):
        vals, ref = sets[i]
        dev = vals.run(rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
    st = sets[2][0].run(rin.FLAG_BATCHED | rin.FLAG_STAGE_TIMES)
    _compare(st, sets[2][1])
    st.release()
    for vals, _ in sets:
        vals.release()
    g.release()


@pytest.mark.parametrize("which", ["config3", "config3_small", "config4", "config4_shuffled", "config1", "mesh"])
def test_batched_mi_matches_oracle_and_replays(ctx, which):
    """The batched material interface (one captured graph per workspace:
    tables for two and three materials, thread cuts beyond, radix-sort dedup)
    gives the oracle's output bit for bit, on the capturing run and on
    replays with plain-form runs between them."""
    import rin

    if which == "mesh":
        m = np.load(S.REF_TESTS + "/mesh.npz")
        pts, tets = m["pts"], m["tets"].astype(np.uint64)
        types, params = S.load_ref_funcs("8-sphere.json")
        g = ctx.grid_explicit(pts, tets)
        hv = S.evaluate(types, params, pts)
        vals = g.upload(hv)
    else:
        if which == "config3":
            res, fs = W.config3()
        elif which == "config3_small":
            _, fs = W.config3()
            res = 24
        elif which == "config4":
            res, fs = W.config4()
        elif which == "config4_shuffled":
            res, fs = W.config4(shuffle_seed=44)
        else:
            res, fs = W.config1()
        types, params = fs.arrays()
        g = ctx.grid_regular(res)
        vals = g.eval(types, params)
        pts, tets = S.grid(res)
        hv = S.evaluate(types, params, pts)
    ref = S.OracleRun(1, pts, tets, hv)
    for rep in range(3):
        dev = vals.run(rin.FLAG_MI | rin.FLAG_BATCHED)
        label = dev.stage_times()[0][0]
        assert label == "batched path (one graph)", label  # no fallback to the plain form
        _compare(dev, ref)
        dev.release()
        vals.run(rin.FLAG_MI).release()
    vals.release()
    g.release()


@pytest.mark.parametrize("form", ["wave", "thread", "wave-scratch"])
@pytest.mark.parametrize("which", ["config3_small", "config4", "config4_shuffled", "mesh"])
def test_batched_mi_cut_forms(ctx, which, form, monkeypatch):
    """Every material-interface tet (tables off) through the lane-group cut or
    the thread cut gives the oracle's output bit for bit, and the records of
    every active tet equal the oracle's.  The lane-group cut of four or more
    materials starts from the 3-material table's complex after three
    materials, or ("-scratch") inserts every material itself."""
    import rin

    if form.endswith("-scratch"):
        monkeypatch.setenv("RIN_MI3_START", "0")
        form = form[: -len("-scratch")]
    monkeypatch.setenv("RIN_CUT_FORM", form)
    if which == "mesh":
        m = np.load(S.REF_TESTS + "/mesh.npz")
        pts, tets = m["pts"], m["tets"].astype(np.uint64)
        types, params = S.load_ref_funcs("8-sphere.json")
        g = ctx.grid_explicit(pts, tets)
        hv = S.evaluate(types, params, pts)
        vals = g.upload(hv)
    else:
        if which == "config3_small":
            _, fs = W.config3()
            res = 24
        elif which == "config4":
            res, fs = W.config4()
        else:
            res, fs = W.config4(shuffle_seed=44)
        types, params = fs.arrays()
        g = ctx.grid_regular(res)
        vals = g.eval(types, params)
        pts, tets = S.grid(res)
        hv = S.evaluate(types, params, pts)
    ref = S.OracleRun(1, pts, tets, hv)
    for rep in range(2):
        dev = vals.run(rin.FLAG_MI | rin.FLAG_BATCHED | rin.FLAG_KEEP_RECORDS)
        label = dev.stage_times()[0][0]
        assert label == "batched path (one graph)", label
        _compare(dev, ref)
        cap = len(ref.words) + 16
        st, w = dev.records(ref.active, cap)
        np.testing.assert_array_equal(st, ref.rec_start)
        np.testing.assert_array_equal(w, ref.words)
        dev.release()
    vals.release()
    g.release()


def test_batched_mi_value_sets_share_the_plan(ctx):
    """Material sets of the same size on one grid share the batched workspace
    and graph (re-recorded and updated in place on each switch)."""
    import rin

    res = 40
    g = ctx.grid_regular(res)
    pts, tets = S.grid(res)
    sets = []
    for seed in (61, 62, 63):
        rng = np.random.default_rng(seed)
        fs = W.FuncSet()
        for _ in range(7):
            fs.sphere(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.0, 0.2))
        types, params = fs.arrays()
        sets.append((g.eval(types, params), S.OracleRun(1, pts, tets, S.evaluate(types, params, pts))))
    for i in (0, 1, 2, 0, 2, 1):
        vals, ref = sets[i]
        dev = vals.run(rin.FLAG_MI | rin.FLAG_BATCHED)
        _compare(dev, ref)
        dev.release()
    for vals, _ in sets:
        vals.release()
    g.release()
