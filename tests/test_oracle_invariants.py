"""Pins the oracle's per-tet cut (the restatement of the external
simplicial_arrangement library called at src/implicit_arrangement.cpp:283 and
src/material_interface.cpp:327) by exact rational geometry, independently of
how the cut is computed:

* brute force: for generic planes / materials, the vertex set is exactly the
  set of label triples (IA) or quadruples (MI) whose intersection point lies
  in the closed tet (and, for MI, where those materials are the maximum);
* every vertex solves to a unique point of the closed tet, and no two
  vertices share a point (also for degenerate integer inputs);
* Euler characteristic V - E + F - C = 1 (the complex is a ball);
* the cells' signed volumes are positive and sum to the tet's volume, which
  fails for any face with a wrong orientation or a wrong cell side;
* every face normal points to its positive side (IA: the supporting plane's
  positive side; MI: the positive label's side), and for MI every cell's
  label is the argmax material at its centroid.

The tet is embedded as corners (0,0,0), (1,0,0), (0,1,0), (0,0,1) (positive
orientation; corner c has barycentric coordinate c)."""
from fractions import Fraction as Q
from itertools import combinations

import numpy as np
import pytest

import support as S

CORNERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
TET_VOL = Q(1, 6)


def _solve(A, b):
    """Exact solution of the square system A x = b, or None when singular."""
    n = len(A)
    M = [list(map(Q, A[i])) + [Q(b[i])] for i in range(n)]
    for c in range(n):
        p = next((r for r in range(c, n) if M[r][c] != 0), None)
        if p is None:
            return None
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def _rows_q(rows):
    return [[Q(float(x)) for x in r] for r in rows]


def _label_row(P, t):
    """Linear form (over barycentric coordinates) of label t: tet face t < 4
    is lambda_t, implicit plane t >= 4 is row t - 4."""
    if t < 4:
        return [Q(int(c == t)) for c in range(4)]
    return P[t - 4]


def _ia_point(P, tri):
    A = [_label_row(P, t) for t in tri] + [[Q(1)] * 4]
    return _solve(A, [0, 0, 0, 1])


def _mi_point(M, quad):
    faces = [t for t in quad if t < 4]
    mats = [t for t in quad if t >= 4]
    A = [[Q(int(c == t)) for c in range(4)] for t in faces]
    for m in mats[1:]:
        A.append([x - y for x, y in zip(M[m - 4], M[mats[0] - 4])])
    A.append([Q(1)] * 4)
    if len(A) != 4:
        return None
    return _solve(A, [0] * (len(A) - 1) + [1])


def _xyz(lam):
    return tuple(sum(lam[c] * CORNERS[c][d] for c in range(4)) for d in range(3))


def _val(row, lam):
    return sum(r * l for r, l in zip(row, lam))


def _sub(a, b):
    return tuple(x - y for x, y in zip(a, b))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def _normal(pts):
    """Newell normal of a planar polygon (exact)."""
    n = [Q(0)] * 3
    m = len(pts)
    for i in range(m):
        a, b = pts[i], pts[(i + 1) % m]
        n[0] += (a[1] - b[1]) * (a[2] + b[2])
        n[1] += (a[2] - b[2]) * (a[0] + b[0])
        n[2] += (a[0] - b[0]) * (a[1] + b[1])
    return tuple(n)


def _grad(row):
    # f = sum_c row[c] * lambda_c with lambda_0 = 1 - x - y - z
    return (row[1] - row[0], row[2] - row[0], row[3] - row[0])


def _cone_volume(pts):
    """Signed volume of the cone from the origin over the polygon (fan)."""
    v = Q(0)
    for i in range(1, len(pts) - 1):
        v += _dot(pts[0], _cross(pts[i], pts[i + 1]))
    return v / 6


def _edges(faces):
    es = set()
    for f in faces:
        vs = f["vs"]
        for i in range(len(vs)):
            a, b = vs[i], vs[(i + 1) % len(vs)]
            es.add((min(a, b), max(a, b)))
    return es


def _check_complex(rec, points, cell_of_side):
    """Shared checks: unique points, Euler, positive cell volumes summing to
    the tet's.  cell_of_side(face) -> (cell on the normal's side, other cell
    or None)."""
    V, F = len(rec["verts"]), len(rec["faces"])
    assert len(set(points)) == V, "two vertices share a point"
    E = len(_edges(rec["faces"]))
    C = rec["ncells"]
    assert V - E + F - C == 1, ("Euler", V, E, F, C)
    vol = {}
    for f in rec["faces"]:
        poly = [points[v] for v in f["vs"]]
        cv = _cone_volume(poly)
        front, back = cell_of_side(f)
        # outward normal of `back` is the cycle normal; of `front`, its reverse
        if back is not None:
            vol[back] = vol.get(back, 0) + cv
        vol[front] = vol.get(front, 0) - cv
    assert len(vol) == C
    assert all(v > 0 for v in vol.values()), vol
    assert sum(vol.values()) == TET_VOL


def _ia_check(rows, generic):
    rec = S.decode_record(S.oracle_cut(0, rows))
    P = _rows_q(rows)
    k = len(rows)
    L = 4 + k
    pts_lam = []
    for v in rec["verts"]:
        tri = v[:3]
        lam = _ia_point(P, tri)
        assert lam is not None, ("vertex is not a point", tri)
        assert all(x >= 0 for x in lam), ("vertex outside the tet", tri)
        pts_lam.append(lam)
    points = [_xyz(l) for l in pts_lam]
    if generic:
        brute = set()
        for tri in combinations(range(L), 3):
            lam = _ia_point(P, tri)
            if lam is not None and all(x >= 0 for x in lam):
                brute.add(tri)
        assert brute == {tuple(v[:3]) for v in rec["verts"]}
    for f in rec["faces"]:
        a = f["a"]
        n = _normal([points[v] for v in f["vs"]])
        assert _dot(n, _grad(_label_row(P, a))) > 0, ("face normal not on the positive side", a)
        for v in f["vs"]:
            assert _val(_label_row(P, a), pts_lam[v]) == 0, "face vertex off its plane"

    def side(f):
        return f["b"], (None if f["c"] == 0xFFFF else f["c"])

    _check_complex(rec, points, side)
    return rec


def _mi_check(rows, generic):
    rec = S.decode_record(S.oracle_cut(1, rows))
    M = _rows_q(rows)
    k = len(rows)
    L = 4 + k
    pts_lam = []
    for v in rec["verts"]:
        lam = _mi_point(M, v)
        assert lam is not None, ("vertex is not a point", v)
        assert all(x >= 0 for x in lam), ("vertex outside the tet", v)
        vals = [_val(r, lam) for r in M]
        top = max(vals)
        assert all(vals[t - 4] == top for t in v if t >= 4), ("vertex materials not maximal", v)
        pts_lam.append(lam)
    points = [_xyz(l) for l in pts_lam]
    if generic:
        brute = set()
        for quad in combinations(range(L), 4):
            if all(t < 4 for t in quad):
                continue
            lam = _mi_point(M, quad)
            if lam is None or any(x < 0 for x in lam):
                continue
            vals = [_val(r, lam) for r in M]
            top = max(vals)
            if all(vals[t - 4] == top for t in quad if t >= 4):
                brute.add(quad)
        assert brute == {tuple(v) for v in rec["verts"]}
    cell_of_label = {lab: c for c, lab in enumerate(rec["cells"])}
    assert len(cell_of_label) == rec["ncells"], "a material with two cells"
    for f in rec["faces"]:
        pos, neg = f["a"], f["b"]
        n = _normal([points[v] for v in f["vs"]])
        if pos < 4:
            g = tuple(-x for x in _grad(_label_row(M, pos)))  # outward across tet face `pos`
        else:
            g = _grad([x - y for x, y in zip(M[pos - 4], M[neg - 4])])
        assert _dot(n, g) > 0, ("face normal not toward the positive label", pos, neg)
    # every cell's label is the argmax at its centroid
    for c, lab in enumerate(rec["cells"]):
        vs = set()
        for f in rec["faces"]:
            if f["b"] == lab or f["a"] == lab:
                vs.update(f["vs"])
        cen = [sum(pts_lam[v][i] for v in vs) / len(vs) for i in range(4)]
        vals = [_val(r, cen) for r in M]
        assert vals[lab - 4] == max(vals), ("cell label is not the maximum", lab)

    _check_mi_volumes(rec, points, cell_of_label)
    return rec


def _check_mi_volumes(rec, points, cell_of_label):
    V, F = len(rec["verts"]), len(rec["faces"])
    assert len(set(points)) == V, "two vertices share a point"
    E = len(_edges(rec["faces"]))
    C = rec["ncells"]
    assert V - E + F - C == 1, ("Euler", V, E, F, C)
    vol = {}
    for f in rec["faces"]:
        cv = _cone_volume([points[v] for v in f["vs"]])
        back = cell_of_label[f["b"]]  # the normal points away from the negative label
        vol[back] = vol.get(back, 0) + cv
        if f["a"] >= 4:
            front = cell_of_label[f["a"]]
            vol[front] = vol.get(front, 0) - cv
    assert len(vol) == C
    assert all(v > 0 for v in vol.values()), vol
    assert sum(vol.values()) == TET_VOL


def _generic_rows(rng, k):
    return rng.standard_normal((k, 4))


def _degenerate_rows(rng, k):
    rows = rng.integers(-2, 3, size=(k, 4)).astype(np.float64)
    if k >= 2 and rng.random() < 0.3:
        rows[-1] = rows[0] * rng.choice([1.0, 2.0, -1.0])
    return rows


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_ia_cut_generic_matches_brute_force(k):
    rng = np.random.default_rng(700 + k)
    for _ in range(40 if k < 4 else 15):
        _ia_check(_generic_rows(rng, k), generic=True)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_ia_cut_degenerate_is_a_valid_complex(k):
    rng = np.random.default_rng(710 + k)
    done = 0
    for _ in range(60):
        rows = _degenerate_rows(rng, k)
        if not np.any(rows):
            continue
        try:
            _ia_check(rows, generic=False)
        except RuntimeError:
            continue  # the oracle refuses the all-zero function
        done += 1
    assert done > 30


@pytest.mark.parametrize("k", [2, 3, 4])
def test_mi_cut_generic_matches_brute_force(k):
    rng = np.random.default_rng(720 + k)
    for _ in range(40 if k < 4 else 15):
        _mi_check(_generic_rows(rng, k), generic=True)


@pytest.mark.parametrize("k", [2, 3])
def test_mi_cut_degenerate_is_a_valid_complex(k):
    rng = np.random.default_rng(730 + k)
    done = 0
    for _ in range(60):
        rows = _degenerate_rows(rng, k)
        try:
            _mi_check(rows, generic=False)
        except RuntimeError:
            continue
        done += 1
    assert done > 30


def test_invariants_catch_a_flipped_face():
    """A face cycle reversed in the record breaks the orientation check."""
    rng = np.random.default_rng(740)
    rows = _generic_rows(rng, 2)
    w = S.oracle_cut(0, rows)
    rec = S.decode_record(w)
    P = _rows_q(rows)
    pts = [_xyz(_ia_point(P, v[:3])) for v in rec["verts"]]
    f = next(f for f in rec["faces"] if f["a"] >= 4)
    n = _normal([pts[v] for v in reversed(f["vs"])])
    assert _dot(n, _grad(P[f["a"] - 4])) < 0
