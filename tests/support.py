"""Test-side helpers: function descriptors, the CPU oracle and the known-answer
glue, all loaded with ctypes from the in-tree builds."""
import ctypes
import json
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_TESTS = os.path.join(GOLDEN, "ref_tests")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u16p = ctypes.POINTER(ctypes.c_uint16)

# function type codes shared by the oracle and the product
PLANE, SPHERE, CYLINDER, TORUS, ZERO = 0, 1, 2, 3, 4


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _unit(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / n, v[1] / n, v[2] / n]


def funcs_from_json(items):
    """Function list (the reference's function-file JSON) -> (types, params[F,12])."""
    F = len(items)
    types = np.zeros(F, dtype=np.int32)
    params = np.zeros((F, 12), dtype=np.float64)
    for j, it in enumerate(items):
        t = it["type"]
        if t == "plane":
            types[j] = PLANE
            params[j, 0:3] = it["point"]
            params[j, 3:6] = it["normal"]
        elif t == "sphere":
            types[j] = SPHERE
            params[j, 0:3] = it["center"]
            params[j, 3] = it["radius"]
            params[j, 4] = 1.0 if it.get("squared", False) else 0.0
        elif t == "cylinder":
            types[j] = CYLINDER
            params[j, 0:3] = it["axis_point"]
            params[j, 3:6] = _unit(it["axis_vector"])
            params[j, 6] = it["radius"]
            params[j, 7] = 1.0 if it.get("is_flipped", False) else 0.0
        elif t == "torus":
            types[j] = TORUS
            params[j, 0:3] = it["center"]
            params[j, 3:6] = _unit(it["axis_vector"])
            params[j, 6] = it["major_radius"]
            params[j, 7] = it["minor_radius"]
            params[j, 8] = 1.0 if it.get("is_flipped", False) else 0.0
        elif t == "zero":
            types[j] = ZERO
        else:
            raise ValueError("unsupported function type " + t)
    return types, params


def load_ref_funcs(name):
    with open(os.path.join(REF_TESTS, name)) as f:
        return funcs_from_json(json.load(f))


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        lib.orc_run.restype = ctypes.c_void_p
        lib.orc_run.argtypes = [ctypes.c_int, _f64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f64p,
                                ctypes.c_uint64, ctypes.c_int]
        lib.orc_free.argtypes = [ctypes.c_void_p]
        lib.orc_sizes.argtypes = [ctypes.c_void_p, _u64p]
        lib.orc_copy_csr.argtypes = [ctypes.c_void_p, _u64p, _u64p]
        lib.orc_copy_verts.argtypes = [ctypes.c_void_p, _u64p, _f64p]
        lib.orc_copy_faces.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _u64p, _u64p]
        lib.orc_copy_records.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u16p]
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_exact_fallbacks.restype = ctypes.c_uint64
        lib.orc_cut.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _u16p, ctypes.c_int]
        lib.orc_timings.argtypes = [ctypes.c_void_p, _f64p, ctypes.c_int]
        lib.orc_timing_label.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_timing_label.restype = ctypes.c_char_p
        _oracle = lib
    return _oracle


def grid(res, bmin=(-1.0, -1.0, -1.0), bmax=(1.0, 1.0, 1.0)):
    N = res + 1
    pts = np.zeros((N ** 3, 3), dtype=np.float64)
    tets = np.zeros((5 * res ** 3, 4), dtype=np.uint64)
    lo = np.asarray(bmin, dtype=np.float64)
    hi = np.asarray(bmax, dtype=np.float64)
    oracle().orc_grid(ctypes.c_int(res), _ptr(lo, _f64p), _ptr(hi, _f64p), _ptr(pts, _f64p), _ptr(tets, _u64p))
    return pts, tets


def evaluate(types, params, pts):
    V = pts.shape[0]
    F = types.shape[0]
    vals = np.zeros((V, F), dtype=np.float64)
    types = np.ascontiguousarray(types, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64)
    rc = oracle().orc_eval(_ptr(types, _i32p), _ptr(params, _f64p), ctypes.c_int(F),
                           _ptr(np.ascontiguousarray(pts), _f64p), ctypes.c_uint64(V), _ptr(vals, _f64p))
    assert rc == 0
    return vals


class OracleRun:
    """Result of the oracle's hot path on one input."""

    def __init__(self, mi, pts, tets, vals, flags=0):
        lib = oracle()
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        h = lib.orc_run(int(mi), _ptr(pts, _f64p), pts.shape[0], _ptr(tets, _u64p), tets.shape[0],
                        _ptr(vals, _f64p), vals.shape[1], flags)
        if not h:
            raise RuntimeError(lib.orc_last_error().decode())
        try:
            s = np.zeros(12, dtype=np.uint64)
            lib.orc_sizes(h, _ptr(s, _u64p))
            self.num_degenerate, self.num_intersecting = int(s[0]), int(s[1])
            self.buckets = [int(s[2]), int(s[3]), int(s[4])]
            nfit, nv, nf, nfv, ntf, nact, nw = (int(x) for x in s[5:12])
            T = tets.shape[0]
            self.func_in_tet = np.zeros(nfit, dtype=np.uint64)
            self.start_index = np.zeros(T + 1, dtype=np.uint64)
            lib.orc_copy_csr(h, _ptr(self.func_in_tet, _u64p), _ptr(self.start_index, _u64p))
            self.verts = np.zeros((nv, 11), dtype=np.uint64)
            self.xyz = np.zeros((nv, 3), dtype=np.float64)
            lib.orc_copy_verts(h, _ptr(self.verts, _u64p), _ptr(self.xyz, _f64p))
            self.face_start = np.zeros(nf + 1, dtype=np.uint64)
            self.face_verts = np.zeros(nfv, dtype=np.uint64)
            self.tf_start = np.zeros(nf + 1, dtype=np.uint64)
            self.tf = np.zeros((ntf, 2), dtype=np.uint64)
            self.face_func = np.zeros((nf, 2), dtype=np.uint64)
            lib.orc_copy_faces(h, _ptr(self.face_start, _u64p), _ptr(self.face_verts, _u64p),
                               _ptr(self.tf_start, _u64p), _ptr(self.tf, _u64p), _ptr(self.face_func, _u64p))
            self.active = np.zeros(nact, dtype=np.uint64)
            self.rec_start = np.zeros(nact + 1, dtype=np.uint64)
            self.words = np.zeros(nw, dtype=np.uint16)
            lib.orc_copy_records(h, _ptr(self.active, _u64p), _ptr(self.rec_start, _u64p), _ptr(self.words, _u16p))
            tim = np.zeros(32, dtype=np.float64)
            n = lib.orc_timings(h, _ptr(tim, _f64p), 32)
            self.timings = {lib.orc_timing_label(h, i).decode(): float(tim[i]) for i in range(n)}
        finally:
            lib.orc_free(h)


def oracle_cut(mi, rows):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros(65535, dtype=np.uint16)
    n = oracle().orc_cut(int(mi), _ptr(rows, _f64p), rows.shape[0], _ptr(out, _u16p), 65535)
    if n < 0:
        raise RuntimeError(oracle().orc_last_error().decode())
    return out[:n].copy()


def decode_record(w):
    """Decode one per-tet cut record (layout of include/rin/cut_record.h)."""
    w = [int(x) for x in w]
    kind, L, nv, nf, nc, flags, length = w[0:7]
    p = 8
    verts = []
    for _ in range(nv):
        verts.append(tuple(w[p:p + 4]))
        p += 4
    faces = []
    for _ in range(nf):
        a, b, c, n = w[p:p + 4]
        p += 4
        faces.append({"a": a, "b": b, "c": c, "vs": w[p:p + n]})
        p += n
    cells = []
    if kind == 2:
        cells = w[p:p + nc]
        p += nc
    groups = orient = None
    if flags & 1:
        groups = w[p:p + L]
        p += L
        if kind == 1:
            orient = w[p:p + L]
            p += L
    assert p == length, (p, length)
    return {"kind": kind, "nlabels": L, "verts": verts, "faces": faces, "ncells": nc, "cells": cells,
            "groups": groups, "orient": orient}


_ka = None


def known_answer_run(mi, pts, tets, vals):
    global _ka
    if _ka is None:
        _ka = ctypes.CDLL(os.path.join(ROOT, "tests", "native", "libka.so"))
        _ka.ka_run.restype = ctypes.c_char_p
        _ka.ka_run.argtypes = [ctypes.c_int, _f64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f64p, ctypes.c_uint64]
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    out = _ka.ka_run(int(mi), _ptr(pts, _f64p), pts.shape[0], _ptr(tets, _u64p), tets.shape[0],
                     _ptr(vals, _f64p), vals.shape[1])
    return json.loads(out.decode())


_dropin = None


def dropin_run(mi, pts, tets, vals, robust=False):
    """The drop-in implicit_arrangement() / material_interface() (GPU path and
    host graph stages) through tests/native/libdropin.so."""
    global _dropin
    if _dropin is None:
        _dropin = ctypes.CDLL(os.path.join(ROOT, "tests", "native", "libdropin.so"))
        _dropin.dropin_run.restype = ctypes.c_char_p
        _dropin.dropin_run.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                       _f64p, ctypes.c_uint64]
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    out = _dropin.dropin_run(int(mi), int(robust), _ptr(pts, _f64p), pts.shape[0], _ptr(tets, _u64p), tets.shape[0],
                             _ptr(vals, _f64p), vals.shape[1])
    return json.loads(out.decode())
