"""Seeded synthetic workloads: regular grids on [-1,1]^3 and function sets of
mixed type (the reference's function-file JSON types plane / sphere /
cylinder / torus).  Seeds are fixed per workload so every run sees the same
input; numpy's PCG64 generator draws every parameter.

Function descriptors are (types[F], params[F,12]) in the layout of
include/rin.h (rin_func_desc).
"""
import math

import numpy as np

PLANE, SPHERE, CYLINDER, TORUS, ZERO = 0, 1, 2, 3, 4


def _unit(v):
    v = [float(x) for x in v]
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / n, v[1] / n, v[2] / n]


class FuncSet:
    def __init__(self):
        self.types = []
        self.params = []
        self.json = []

    def plane(self, point, normal):
        p = np.zeros(12)
        p[0:3] = point
        p[3:6] = normal
        self.types.append(PLANE)
        self.params.append(p)
        self.json.append({"type": "plane", "point": [float(x) for x in point], "normal": [float(x) for x in normal]})

    def sphere(self, center, radius, squared=False):
        p = np.zeros(12)
        p[0:3] = center
        p[3] = radius
        p[4] = 1.0 if squared else 0.0
        self.types.append(SPHERE)
        self.params.append(p)
        self.json.append({"type": "sphere", "center": [float(x) for x in center], "radius": float(radius),
                          "squared": bool(squared)})

    def cylinder(self, point, axis, radius, flipped=False):
        a = _unit(axis)
        p = np.zeros(12)
        p[0:3] = point
        p[3:6] = a
        p[6] = radius
        p[7] = 1.0 if flipped else 0.0
        self.types.append(CYLINDER)
        self.params.append(p)
        self.json.append({"type": "cylinder", "axis_point": [float(x) for x in point], "axis_vector": a,
                          "radius": float(radius), "is_flipped": bool(flipped)})

    def torus(self, center, axis, major, minor, flipped=False):
        a = _unit(axis)
        p = np.zeros(12)
        p[0:3] = center
        p[3:6] = a
        p[6] = major
        p[7] = minor
        p[8] = 1.0 if flipped else 0.0
        self.types.append(TORUS)
        self.params.append(p)
        self.json.append({"type": "torus", "center": [float(x) for x in center], "axis_vector": a,
                          "major_radius": float(major), "minor_radius": float(minor), "is_flipped": bool(flipped)})

    def copy_of(self, j):
        self.types.append(self.types[j])
        self.params.append(self.params[j].copy())
        self.json.append(dict(self.json[j]))

    def arrays(self):
        return np.asarray(self.types, dtype=np.int32), np.asarray(self.params, dtype=np.float64).reshape(-1, 12)

    def permuted(self, order):
        out = FuncSet()
        out.types = [self.types[i] for i in order]
        out.params = [self.params[i] for i in order]
        out.json = [self.json[i] for i in order]
        return out

    def __len__(self):
        return len(self.types)


def _rand_unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def mixed(rng, n_plane, n_sphere, n_cyl, n_torus):
    fs = FuncSet()
    for _ in range(n_plane):
        n = _rand_unit(rng)
        d = rng.uniform(-0.5, 0.5)
        fs.plane(d * n, n)
    for _ in range(n_sphere):
        fs.sphere(rng.uniform(-0.5, 0.5, 3), rng.uniform(0.2, 0.6))
    for _ in range(n_cyl):
        fs.cylinder(rng.uniform(-0.5, 0.5, 3), _rand_unit(rng), rng.uniform(0.1, 0.4))
    for _ in range(n_torus):
        fs.torus(rng.uniform(-0.5, 0.5, 3), _rand_unit(rng), rng.uniform(0.3, 0.6), rng.uniform(0.05, 0.2))
    return fs


def config1():
    """res 20, 3 planes + 2 spheres (implicit arrangement, CPU-sized)."""
    return 20, mixed(np.random.default_rng(1), 3, 2, 0, 0)


def config2():
    """res 100 (101^3 vertices, 5,000,000 tets), 6 planes, 6 spheres, 4 cylinders, 4 tori."""
    return 100, mixed(np.random.default_rng(2), 6, 6, 4, 4)


def config3():
    """material interface, res 64, 8 materials -|x - p_i| + w_i (Voronoi-like)."""
    rng = np.random.default_rng(3)
    fs = FuncSet()
    for _ in range(8):
        fs.sphere(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.0, 0.2))
    return 64, fs


def config4(shuffle_seed=None):
    """dyadic res 32 (spacing 1/16): 12 axis-aligned planes on grid coordinates,
    4 planes through exactly three grid vertices with integer-scaled normals,
    and 2 duplicated functions; optionally shuffled."""
    rng = np.random.default_rng(4)
    fs = FuncSet()
    for j in range(12):
        axis = j % 3
        c = int(rng.integers(-14, 15)) / 16.0
        pt = [0.0, 0.0, 0.0]
        pt[axis] = c
        nrm = [0.0, 0.0, 0.0]
        nrm[axis] = 1.0
        fs.plane(pt, nrm)
    made = 0
    while made < 4:
        P = rng.integers(-16, 17, size=(3, 3)) / 16.0
        n = np.cross(P[1] - P[0], P[2] - P[0]) * 256.0  # integer components
        if not np.any(n):
            continue
        fs.plane(P[0], n)
        made += 1
    fs.copy_of(0)
    fs.copy_of(13)
    if shuffle_seed is not None:
        order = np.random.default_rng(shuffle_seed).permutation(len(fs))
        fs = fs.permuted(order)
    return 32, fs


def serving_job(i):
    """one serving request: grid res in {16, 32, 64}, 2..32 functions of mixed
    type, 70% implicit arrangement / 30% material interface (seed 5e6 + i)."""
    rng = np.random.default_rng(5_000_000 + i)
    res = int(rng.choice([16, 32, 64]))
    F = int(rng.integers(2, 33))
    mi = bool(rng.random() < 0.3)
    if mi:
        fs = FuncSet()
        for _ in range(F):
            fs.sphere(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.0, 0.2))
    else:
        counts = rng.multinomial(F, [0.3, 0.3, 0.2, 0.2])
        fs = mixed(rng, *[int(c) for c in counts])
    return res, fs, mi
