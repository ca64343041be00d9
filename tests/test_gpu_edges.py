"""Iso-edges, patches and chains on the device (rin_result_edges,
rin_result_patches, rin_result_chains) against compute_mesh_edges,
compute_patches and compute_chains (src/mesh_connectivity.cpp:10-93, 196-241)
restated over the oracle's faces: edge ids in order of first occurrence over
(face, j), each edge's endpoints (smaller first), its (face, j) list in that
order, and each face's edge list; patches seeded at their smallest unvisited
face, faces in queue order across manifold edges; chains seeded at their
smallest unvisited non-manifold edge, edges in queue order across vertices
with exactly two non-manifold edges."""
from collections import deque

import numpy as np
import pytest

import support as S
import workloads as W

pytestmark = pytest.mark.gpu


def _edges_of_faces(face_start, face_verts):
    """The reference's loop: face edge j joins vertices j and j + 1
    (cyclically); a new sorted pair gets the next id."""
    ids, v12, occ, eof = {}, [], [], []
    for f in range(len(face_start) - 1):
        b, e = int(face_start[f]), int(face_start[f + 1])
        n = e - b
        for j in range(n):
            a, c = int(face_verts[b + j]), int(face_verts[b + (j + 1) % n])
            k = (a, c) if a < c else (c, a)
            i = ids.setdefault(k, len(v12))
            if i == len(v12):
                v12.append(k)
                occ.append([])
            occ[i].append((f, j))
            eof.append(i)
    return v12, occ, eof


def _check(dev, ref):
    v12, st, occ, eo = dev.edges()
    rv, rocc, reof = _edges_of_faces(ref.face_start, ref.face_verts)
    assert len(v12) == len(rv)
    np.testing.assert_array_equal(v12, np.array(rv, dtype=np.uint64).reshape(-1, 2))
    np.testing.assert_array_equal(np.diff(st.astype(np.int64)), [len(o) for o in rocc])
    flat = np.array([p for o in rocc for p in o], dtype=np.uint64).reshape(-1, 2)
    np.testing.assert_array_equal(occ, flat)
    np.testing.assert_array_equal(eo, np.array(reof, dtype=np.uint64))


@pytest.mark.parametrize("mi,which", [(0, "config1"), (0, "config4"), (1, "config4"), (1, "config3_small")])
def test_edges_match_reference_loop(ctx, mi, which):
    import rin

    if which == "config1":
        res, fs = W.config1()
    elif which == "config4":
        res, fs = W.config4(shuffle_seed=44)
    else:
        _, fs = W.config3()
        res = 24
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(rin.FLAG_MI if mi else 0)
    pts, tets = S.grid(res)
    ref = S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))
    _check(dev, ref)
    # cached: a second request returns the same lists
    a = dev.edges()
    b = dev.edges()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    dev.release()
    vals.release()
    g.release()


def test_edges_empty(ctx):
    """No iso-faces: no edges."""
    g = ctx.grid_regular(8)
    fs = W.mixed(np.random.default_rng(3), 0, 0, 0, 0)
    fs.plane([0.0, 0.0, 5.0], [0.0, 0.0, 1.0])
    types, params = fs.arrays()
    vals = g.eval(types, params)
    dev = vals.run(0)
    assert dev.sizes.n_faces == 0
    v12, st, occ, eo = dev.edges()
    assert v12.shape == (0, 2) and list(st) == [0] and occ.shape == (0, 2) and eo.shape == (0,)


def test_edges_headline_properties(ctx):
    """At the 101^3-vertex, 20-function size: every occurrence's face edge is its
    edge's vertex pair, each edge's occurrences are listed in (face, j) order,
    every face-edge position is listed once, and edge ids follow first
    occurrence (the first occurrences increase with the id)."""
    res, fs = W.config2()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(0)
    vs, v, _, _, _ = dev.faces()
    v12, st, occ, eo = dev.edges()
    ne, nfv = len(v12), len(v)
    assert ne > 0 and st[-1] == nfv
    f, j = occ[:, 0].astype(np.int64), occ[:, 1].astype(np.int64)
    pos = vs[f].astype(np.int64) + j
    size = (vs[f + 1] - vs[f]).astype(np.int64)
    assert np.all(j < size)
    nxt = vs[f].astype(np.int64) + (j + 1) % size
    a, b = v[pos], v[nxt]
    eid = np.repeat(np.arange(ne), np.diff(st.astype(np.int64)))
    np.testing.assert_array_equal(np.minimum(a, b), v12[eid, 0])
    np.testing.assert_array_equal(np.maximum(a, b), v12[eid, 1])
    assert np.all(v12[:, 0] < v12[:, 1])
    # every face-edge position once, and edge_of agrees with the lists
    np.testing.assert_array_equal(np.sort(pos), np.arange(nfv))
    np.testing.assert_array_equal(eo[pos], eid)
    # occurrence order within an edge, and first-occurrence order of ids
    same = eid[1:] == eid[:-1]
    assert np.all(pos[1:][same] > pos[:-1][same])
    firsts = pos[st[:-1].astype(np.int64)]
    assert np.all(np.diff(firsts) > 0)
    # distinct pairs: no two ids share a key
    key = v12[:, 0] * (1 << 32) + v12[:, 1]
    assert len(np.unique(key)) == ne
    dev.release()
    vals.release()
    g.release()


def _patches_of(face_start, face_verts):
    """The reference's loop: a queue from each unvisited face, neighbours
    across edges with exactly two occurrences, in the face's edge order."""
    v12, occ, eof = _edges_of_faces(face_start, face_verts)
    nf = len(face_start) - 1
    seen = [False] * nf
    patches = []
    k = 0
    face_edges = []
    for f in range(nf):
        n = int(face_start[f + 1] - face_start[f])
        face_edges.append(eof[k:k + n])
        k += n
    for i in range(nf):
        if seen[i]:
            continue
        patch = [i]
        seen[i] = True
        q = deque([i])
        while q:
            f = q.popleft()
            for e in face_edges[f]:
                if len(occ[e]) != 2:
                    continue
                g = occ[e][1][0] if occ[e][0][0] == f else occ[e][0][0]
                if not seen[g]:
                    seen[g] = True
                    q.append(g)
                    patch.append(g)
        patches.append(patch)
    return patches


@pytest.mark.parametrize("mi,which", [(0, "config1"), (0, "config4"), (1, "config4"), (1, "config3_small"),
                                      (0, "config2_res48")])
def test_patches_match_reference_loop(ctx, mi, which):
    import rin

    if which == "config1":
        res, fs = W.config1()
    elif which == "config4":
        res, fs = W.config4(shuffle_seed=44)
    elif which == "config2_res48":
        _, fs = W.config2()
        res = 48
    else:
        _, fs = W.config3()
        res = 24
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    for flags in (0, rin.FLAG_BATCHED):
        dev = vals.run((rin.FLAG_MI if mi else 0) | flags)
        pts, tets = S.grid(res)
        ref = S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))
        want = _patches_of(ref.face_start, ref.face_verts)
        st, fc = dev.patches()
        assert len(st) - 1 == len(want)
        got = [list(fc[int(st[p]):int(st[p + 1])]) for p in range(len(want))]
        assert got == want
        # cached: a second request returns the same lists
        st2, fc2 = dev.patches()
        np.testing.assert_array_equal(st2, st)
        np.testing.assert_array_equal(fc2, fc)
        dev.release()
    vals.release()
    g.release()


def test_patches_empty(ctx):
    g = ctx.grid_regular(8)
    fs = W.mixed(np.random.default_rng(3), 0, 0, 0, 0)
    fs.plane([0.0, 0.0, 5.0], [0.0, 0.0, 1.0])
    types, params = fs.arrays()
    vals = g.eval(types, params)
    dev = vals.run(0)
    st, fc = dev.patches()
    assert list(st) == [0] and len(fc) == 0


def test_patches_headline(ctx):
    """At the 101^3-vertex, 20-function size, the device patches equal the
    reference loop run over the device's own faces."""
    import rin

    res, fs = W.config2()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(rin.FLAG_BATCHED)
    vs, v, _, _, _ = dev.faces()
    want = _patches_of(vs, v)
    st, fc = dev.patches()
    assert len(st) - 1 == len(want)
    flat = np.array([f for p in want for f in p], dtype=np.uint64)
    np.testing.assert_array_equal(fc, flat)
    np.testing.assert_array_equal(np.diff(st.astype(np.int64)), [len(p) for p in want])
    dev.release()
    vals.release()
    g.release()



def _chains_of(v12, nocc):
    """The reference's loop (src/mesh_connectivity.cpp:196-241) over edges with
    endpoints v12 and nocc face occurrences each; the per-vertex lists of
    non-manifold edges are built in edge order, as its caller builds them."""
    nm = {}
    for i, (a, b) in enumerate(v12):
        if nocc[i] > 2:
            nm.setdefault(int(a), []).append(i)
            nm.setdefault(int(b), []).append(i)
    seen = [False] * len(v12)
    chains = []
    for i in range(len(v12)):
        if seen[i] or nocc[i] <= 2:
            continue
        chain = [i]
        seen[i] = True
        q = deque([i])
        while q:
            e = q.popleft()
            for v in (int(v12[e][0]), int(v12[e][1])):
                lst = nm.get(v, [])
                if len(lst) != 2:
                    continue
                o = lst[1] if lst[0] == e else lst[0]
                if not seen[o]:
                    seen[o] = True
                    q.append(o)
                    chain.append(o)
        chains.append(chain)
    return chains


def _check_chains(dev, want):
    st, ed = dev.chains()
    assert len(st) - 1 == len(want)
    flat = np.array([e for c in want for e in c], dtype=np.uint64)
    np.testing.assert_array_equal(ed, flat)
    np.testing.assert_array_equal(np.diff(st.astype(np.int64)), [len(c) for c in want])
    # cached: a second request returns the same lists
    st2, ed2 = dev.chains()
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(ed2, ed)


@pytest.mark.parametrize("mi,which", [(0, "config1"), (0, "config4"), (1, "config4"), (1, "config3_small"),
                                      (0, "config2_res48")])
def test_chains_match_reference_loop(ctx, mi, which):
    """Chains over the oracle's faces (its own edges, restated), for both forms."""
    import rin

    if which == "config1":
        res, fs = W.config1()
    elif which == "config4":
        res, fs = W.config4(shuffle_seed=44)
    elif which == "config2_res48":
        _, fs = W.config2()
        res = 48
    else:
        _, fs = W.config3()
        res = 24
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    pts, tets = S.grid(res)
    ref = S.OracleRun(mi, pts, tets, S.evaluate(types, params, pts))
    rv, rocc, _ = _edges_of_faces(ref.face_start, ref.face_verts)
    want = _chains_of(rv, [len(o) for o in rocc])
    assert len(want) > 0
    for flags in (0, rin.FLAG_BATCHED):
        dev = vals.run((rin.FLAG_MI if mi else 0) | flags)
        _check_chains(dev, want)
        dev.release()
    vals.release()
    g.release()


def test_chains_empty(ctx):
    """One function: no non-manifold edges, no chains; no faces at all: none."""
    g = ctx.grid_regular(8)
    for planes in ([[0.0, 0.0, 0.5]], [[0.0, 0.0, 5.0]]):
        fs = W.mixed(np.random.default_rng(3), 0, 0, 0, 0)
        fs.plane(planes[0], [0.0, 0.0, 1.0])
        types, params = fs.arrays()
        vals = g.eval(types, params)
        dev = vals.run(0)
        st, ed = dev.chains()
        assert list(st) == [0] and len(ed) == 0
        dev.release()
        vals.release()


def test_chains_headline(ctx):
    """At the 101^3-vertex, 20-function size, the device chains equal the
    reference loop run over the device's own edges; the workload has both
    open chains and closed ones (cycles: the last edge meets the first)."""
    import rin

    res, fs = W.config2()
    types, params = fs.arrays()
    g = ctx.grid_regular(res)
    vals = g.eval(types, params)
    dev = vals.run(rin.FLAG_BATCHED)
    v12, st, _, _ = dev.edges()
    want = _chains_of(v12, np.diff(st.astype(np.int64)))
    _check_chains(dev, want)
    dev.release()
    vals.release()
    g.release()
