"""ctypes binding of the C ABI in include/rin.h (librin.so, built in-tree).

This is the Python face of the drop-in boundary: the same entry points a
reference-side FFI would bind.  It fails loudly when the HIP library is not
built; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librin.so")

FLAG_MI = 1
FLAG_REVERSE = 2
FLAG_NO_LOOKUP = 4
FLAG_BATCHED = 8
FLAG_KEEP_RECORDS = 16
FLAG_CUT_ONLY = 32
FLAG_STAGE_TIMES = 64
FLAG_FIXED_FORMS = 128

STATUS = {0: "ok", -1: "bad argument", -2: "device error", -3: "out of device memory", -4: "capacity exceeded",
          -5: "unsupported", -6: "internal error"}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class FuncDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("reserved", ctypes.c_int32), ("params", ctypes.c_double * 12)]


class Sizes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("n_tets", "n_funcs", "num_degenerate", "num_intersecting", "b0", "b1", "b2", "n_func_in_tet",
                 "n_verts", "n_faces", "n_face_verts", "n_tet_faces", "n_vert_occ", "n_face_occ",
                 "n_face_occ_verts", "n_cut34", "n_cut58", "n_rec_words", "cut_k34", "cut_k58", "cut_words34",
                 "cut_words58")]


class RinError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rin error %d (%s): %s" % (code, STATUS.get(code, "?"), msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librin.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.rin_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int32]
        L.rin_destroy.argtypes = [_vp]
        L.rin_last_error.argtypes = [_vp]
        L.rin_last_error.restype = ctypes.c_char_p
        L.rin_grid_regular.argtypes = [_vp, ctypes.c_int32, _f64p, _f64p, ctypes.POINTER(_vp)]
        L.rin_grid_explicit.argtypes = [_vp, _f64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.rin_grid_points.argtypes = [_vp, _vp, _f64p]
        L.rin_grid_release.argtypes = [_vp, _vp]
        L.rin_values_eval.argtypes = [_vp, _vp, ctypes.POINTER(FuncDesc), ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.rin_values_upload.argtypes = [_vp, _vp, _f64p, ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.rin_values_download.argtypes = [_vp, _vp, _f64p]
        L.rin_values_release.argtypes = [_vp, _vp]
        L.rin_run.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.rin_result_sizes.argtypes = [_vp, _vp, ctypes.POINTER(Sizes)]
        L.rin_result_csr.argtypes = [_vp, _vp, _u64p, _u64p]
        L.rin_result_verts.argtypes = [_vp, _vp, _u64p, _f64p]
        L.rin_result_faces.argtypes = [_vp, _vp, _u64p, _u64p, _u64p, _u64p, _u64p]
        L.rin_result_records.argtypes = [_vp, _vp, _u64p, ctypes.c_uint64, _u64p, _u16p, ctypes.c_uint64]
        L.rin_result_cut_counts.argtypes = [_vp, _vp, _u64p]
        L.rin_result_edge_count.argtypes = [_vp, _vp, _u64p]
        L.rin_result_edges.argtypes = [_vp, _vp, _u64p, _u64p, _u64p, _u64p]
        L.rin_result_patch_count.argtypes = [_vp, _vp, _u64p]
        L.rin_result_patches.argtypes = [_vp, _vp, _u64p, _u64p]
        L.rin_result_chain_count.argtypes = [_vp, _vp, _u64p]
        L.rin_result_chains.argtypes = [_vp, _vp, _u64p, _u64p]
        L.rin_result_stage_times.argtypes = [_vp, _vp, _f64p, ctypes.c_int32]
        L.rin_result_stage_label.argtypes = [_vp, _vp, ctypes.c_int32]
        L.rin_result_stage_label.restype = ctypes.c_char_p
        L.rin_result_release.argtypes = [_vp, _vp]
        L.rin_debug_cuts.argtypes = [_vp, ctypes.c_int32, _f64p, _i32p, ctypes.c_int32, ctypes.c_int32, _u16p,
                                     _i32p]
        L.rin_debug_copy.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        L.rin_debug_sqrt.argtypes = [_vp, _f64p, ctypes.c_uint64, _f64p]
        L.rin_debug_scan.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                     _u32p, _u32p]
        L.rin_debug_table_cuts.argtypes = [ctypes.c_int32, _f64p, ctypes.c_int32, ctypes.c_int32, _u16p, _i32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def descs(types, params):
    F = len(types)
    arr = (FuncDesc * F)()
    for j in range(F):
        arr[j].type = int(types[j])
        for q in range(12):
            arr[j].params[q] = float(params[j][q])
    return arr


def debug_table_cuts(mode, rows, cap=512):
    """Host-only lookup of the tables (no device): mode 0 IA one-function
    table, rows (n, 4); mode 1 IA two-function table, rows (n, 2, 4); mode 2
    MI two-material table, rows (n, 2, 4); mode 3 MI three-material table,
    rows (n, 3, 4).  Returns a list of uint16 records, or negative codes where
    the table has no entry."""
    nr = 1 if mode == 0 else (3 if mode == 3 else 2)
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, nr * 4))
    n = rows.shape[0]
    words = np.zeros(n * cap, dtype=np.uint16)
    lens = np.zeros(n, dtype=np.int32)
    rc = lib().rin_debug_table_cuts(mode, _p(rows, _f64p), n, cap, _p(words, _u16p), _p(lens, _i32p))
    if rc != 0:
        raise RinError(rc, "rin_debug_table_cuts failed")
    return [words[i * cap:i * cap + int(lens[i])].copy() if lens[i] > 0 else int(lens[i]) for i in range(n)]


class Context:
    def __init__(self, device=0):
        self.h = _vp()
        rc = lib().rin_create(ctypes.byref(self.h), device)
        if rc != 0:
            raise RinError(rc, "rin_create failed (no GPU?)")

    def check(self, rc):
        if rc != 0:
            raise RinError(rc, lib().rin_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().rin_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # grids
    def grid_regular(self, res, bmin=(-1.0, -1.0, -1.0), bmax=(1.0, 1.0, 1.0)):
        g = _vp()
        lo = np.asarray(bmin, dtype=np.float64)
        hi = np.asarray(bmax, dtype=np.float64)
        self.check(lib().rin_grid_regular(self.h, res, _p(lo, _f64p), _p(hi, _f64p), ctypes.byref(g)))
        N = res + 1
        return Grid(self, g, N ** 3, 5 * res ** 3)

    def grid_explicit(self, pts, tets):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.uint64)
        g = _vp()
        self.check(lib().rin_grid_explicit(self.h, _p(pts, _f64p), pts.shape[0], _p(tets, _u64p), tets.shape[0],
                                           ctypes.byref(g)))
        return Grid(self, g, pts.shape[0], tets.shape[0])

    def debug_scan(self, fields, stride, data, n, count=None):
        """Device exclusive scans of `data` (uint32; layout as rin_debug_scan)."""
        data = np.ascontiguousarray(data, dtype=np.uint32)
        out = np.zeros(fields * n, dtype=np.uint32)
        tot = np.zeros(fields, dtype=np.uint32)
        self.check(lib().rin_debug_scan(self.h, fields, stride, _p(data, _u32p), n, n if count is None else count,
                                        _p(out, _u32p), _p(tot, _u32p)))
        return out.reshape(fields, n), tot

    def debug_sqrt(self, x):
        """The device evaluation's correctly rounded square root of float64 x."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self.check(lib().rin_debug_sqrt(self.h, _p(x, _f64p), x.size, _p(out, _f64p)))
        return out

    def debug_copy(self, nbytes, reps=5):
        """Average milliseconds of one plain device copy of nbytes."""
        ms = ctypes.c_double(0)
        self.check(lib().rin_debug_copy(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def debug_cuts(self, mi, cases, cap=4096):
        """cases: list of (k, 4) arrays -> list of uint16 records (or negative codes)."""
        ks = np.asarray([len(c) for c in cases], dtype=np.int32)
        rows = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 4) for c in cases]))
        n = len(cases)
        words = np.zeros(n * cap, dtype=np.uint16)
        lens = np.zeros(n, dtype=np.int32)
        self.check(lib().rin_debug_cuts(self.h, int(mi), _p(rows, _f64p), _p(ks, _i32p), n, cap, _p(words, _u16p),
                                        _p(lens, _i32p)))
        out = []
        for i in range(n):
            L = int(lens[i])
            out.append(words[i * cap:i * cap + L].copy() if L > 0 else L)
        return out


class Grid:
    def __init__(self, ctx, h, V, T):
        self.ctx, self.h, self.V, self.T = ctx, h, V, T

    def points(self):
        out = np.zeros((self.V, 3), dtype=np.float64)
        self.ctx.check(lib().rin_grid_points(self.ctx.h, self.h, _p(out, _f64p)))
        return out

    def eval(self, types, params):
        d = descs(types, params)
        v = _vp()
        self.ctx.check(lib().rin_values_eval(self.ctx.h, self.h, d, len(types), ctypes.byref(v)))
        return Values(self, v, len(types))

    def upload(self, vals_rowmajor):
        a = np.ascontiguousarray(vals_rowmajor, dtype=np.float64)
        v = _vp()
        self.ctx.check(lib().rin_values_upload(self.ctx.h, self.h, _p(a, _f64p), a.shape[1], ctypes.byref(v)))
        return Values(self, v, a.shape[1])

    def release(self):
        if self.h:
            lib().rin_grid_release(self.ctx.h, self.h)
            self.h = None


class Values:
    def __init__(self, grid, h, F):
        self.grid, self.h, self.F = grid, h, F

    def download(self):
        out = np.zeros((self.grid.V, self.F), dtype=np.float64)
        self.grid.ctx.check(lib().rin_values_download(self.grid.ctx.h, self.h, _p(out, _f64p)))
        return out

    def run(self, flags=0):
        r = _vp()
        ctx = self.grid.ctx
        ctx.check(lib().rin_run(ctx.h, self.grid.h, self.h, flags, ctypes.byref(r)))
        return Result(ctx, r, self.grid.T)

    def release(self):
        if self.h:
            lib().rin_values_release(self.grid.ctx.h, self.h)
            self.h = None


class Result:
    def __init__(self, ctx, h, T):
        self.ctx, self.h, self.T = ctx, h, T
        s = Sizes()
        ctx.check(lib().rin_result_sizes(ctx.h, h, ctypes.byref(s)))
        self.sizes = s
        self.num_degenerate = s.num_degenerate
        self.num_intersecting = s.num_intersecting
        self.buckets = [s.b0, s.b1, s.b2]

    def csr(self):
        s = self.sizes
        fit = np.zeros(s.n_func_in_tet, dtype=np.uint64)
        st = np.zeros(s.n_tets + 1, dtype=np.uint64)
        self.ctx.check(lib().rin_result_csr(self.ctx.h, self.h, _p(fit, _u64p), _p(st, _u64p)))
        return fit, st

    def verts(self):
        n = self.sizes.n_verts
        rec = np.zeros((n, 11), dtype=np.uint64)
        xyz = np.zeros((n, 3), dtype=np.float64)
        self.ctx.check(lib().rin_result_verts(self.ctx.h, self.h, _p(rec, _u64p), _p(xyz, _f64p)))
        return rec, xyz

    def faces(self):
        s = self.sizes
        vs = np.zeros(s.n_faces + 1, dtype=np.uint64)
        v = np.zeros(s.n_face_verts, dtype=np.uint64)
        ts = np.zeros(s.n_faces + 1, dtype=np.uint64)
        tf = np.zeros((s.n_tet_faces, 2), dtype=np.uint64)
        fn = np.zeros((s.n_faces, 2), dtype=np.uint64)
        self.ctx.check(lib().rin_result_faces(self.ctx.h, self.h, _p(vs, _u64p), _p(v, _u64p), _p(ts, _u64p),
                                              _p(tf, _u64p), _p(fn, _u64p)))
        return vs, v, ts, tf, fn

    def records(self, tets, cap=None):
        tets = np.ascontiguousarray(tets, dtype=np.uint64)
        cap = cap or 8192 * max(1, len(tets))
        st = np.zeros(len(tets) + 1, dtype=np.uint64)
        w = np.zeros(cap, dtype=np.uint16)
        self.ctx.check(lib().rin_result_records(self.ctx.h, self.h, _p(tets, _u64p), len(tets), _p(st, _u64p),
                                                _p(w, _u16p), cap))
        return st, w[:int(st[-1])]

    def edges(self):
        """Iso-edges on the device: (v12 (NE, 2), start (NE + 1), occ (NFV, 2) of
        (face, j), edge_of (NFV))."""
        ne = ctypes.c_uint64(0)
        self.ctx.check(lib().rin_result_edge_count(self.ctx.h, self.h, ctypes.byref(ne)))
        nfv = self.sizes.n_face_verts
        v12 = np.zeros((ne.value, 2), dtype=np.uint64)
        st = np.zeros(ne.value + 1, dtype=np.uint64)
        occ = np.zeros((nfv, 2), dtype=np.uint64)
        eo = np.zeros(nfv, dtype=np.uint64)
        self.ctx.check(lib().rin_result_edges(self.ctx.h, self.h, _p(v12, _u64p), _p(st, _u64p), _p(occ, _u64p),
                                              _p(eo, _u64p)))
        return v12, st, occ, eo

    def patches(self):
        """Patches on the device: (start (NP + 1), faces (NF))."""
        n = ctypes.c_uint64(0)
        self.ctx.check(lib().rin_result_patch_count(self.ctx.h, self.h, ctypes.byref(n)))
        st = np.zeros(n.value + 1, dtype=np.uint64)
        fc = np.zeros(max(1, self.sizes.n_faces), dtype=np.uint64)
        self.ctx.check(lib().rin_result_patches(self.ctx.h, self.h, _p(st, _u64p), _p(fc, _u64p)))
        return st, fc[:self.sizes.n_faces]

    def chains(self):
        """Chains of non-manifold edges on the device: (start (NC + 1), edges)."""
        n = ctypes.c_uint64(0)
        self.ctx.check(lib().rin_result_chain_count(self.ctx.h, self.h, ctypes.byref(n)))
        ne = ctypes.c_uint64(0)
        self.ctx.check(lib().rin_result_edge_count(self.ctx.h, self.h, ctypes.byref(ne)))
        st = np.zeros(n.value + 1, dtype=np.uint64)
        ed = np.zeros(max(1, ne.value), dtype=np.uint64)
        self.ctx.check(lib().rin_result_chains(self.ctx.h, self.h, _p(st, _u64p), _p(ed, _u64p)))
        return st, ed[:int(st[-1])]

    def cut_counts(self):
        n = self.sizes.num_intersecting
        c = np.zeros((n, 3), dtype=np.uint64)
        self.ctx.check(lib().rin_result_cut_counts(self.ctx.h, self.h, _p(c, _u64p)))
        return c

    def stage_times(self):
        L = lib()
        ms = np.zeros(64, dtype=np.float64)
        n = L.rin_result_stage_times(self.ctx.h, self.h, _p(ms, _f64p), 64)
        return [(L.rin_result_stage_label(self.ctx.h, self.h, i).decode(), float(ms[i])) for i in range(min(n, 64))]

    def release(self):
        if self.h:
            lib().rin_result_release(self.ctx.h, self.h)
            self.h = None


# ---------------------------------------------------------------- serving loop (include/rin/serve.h)

SERVE_LIB_PATH = os.path.join(HERE, "librin_serve.so")
_serve = None


class JobDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("res", ctypes.c_int32), ("n_funcs", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("funcs", ctypes.POINTER(FuncDesc))]


class JobInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("slot", ctypes.c_int32),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("hbm_estimate", ctypes.c_uint64), ("sizes", Sizes)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("max_in_flight", ctypes.c_uint64), ("tets", ctypes.c_uint64), ("first_submit", ctypes.c_double),
                ("last_end", ctypes.c_double), ("p50_latency", ctypes.c_double), ("p99_latency", ctypes.c_double),
                ("max_latency", ctypes.c_double), ("hbm_budget", ctypes.c_uint64)]


def serve_lib():
    global _serve
    if _serve is None:
        if not os.path.exists(SERVE_LIB_PATH):
            raise RuntimeError("librin_serve.so is not built (make -C %s)" % HERE)
        lib()
        L = ctypes.CDLL(SERVE_LIB_PATH)
        L.rin_server_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.rin_server_destroy.argtypes = [_vp]
        L.rin_server_last_error.argtypes = [_vp]
        L.rin_server_last_error.restype = ctypes.c_char_p
        L.rin_server_submit.argtypes = [_vp, ctypes.POINTER(JobDesc), _u64p]
        L.rin_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(JobInfo)]
        L.rin_server_job_verts.argtypes = [_vp, ctypes.c_uint64, _u64p, _f64p]
        L.rin_server_job_faces.argtypes = [_vp, ctypes.c_uint64, _u64p, _u64p, _u64p, _u64p, _u64p]
        L.rin_server_job_release.argtypes = [_vp, ctypes.c_uint64]
        L.rin_server_stats_get.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.rin_server_now.argtypes = [_vp]
        L.rin_server_now.restype = ctypes.c_double
        _serve = L
    return _serve


class Server:
    """Continuous-batching server on one GPU: n_slots concurrent jobs, FIFO
    admission by free slot and HBM budget (bytes; 0 = 80% of free memory)."""

    def __init__(self, device=0, slots=4, hbm_budget=0):
        self.h = _vp()
        rc = serve_lib().rin_server_create(device, slots, int(hbm_budget), ctypes.byref(self.h))
        if rc != 0:
            raise RinError(rc, "rin_server_create failed (no GPU?)")

    def check(self, rc):
        if rc != 0:
            raise RinError(rc, serve_lib().rin_server_last_error(self.h).decode())

    def now(self):
        return serve_lib().rin_server_now(self.h)

    def submit(self, mi, res, types, params, flags=0):
        arr = descs(types, params)
        d = JobDesc(1 if mi else 0, int(res), len(types), int(flags), arr)
        jid = ctypes.c_uint64(0)
        self.check(serve_lib().rin_server_submit(self.h, ctypes.byref(d), ctypes.byref(jid)))
        return jid.value

    def wait(self, jid, timeout=120.0):
        info = JobInfo()
        self.check(serve_lib().rin_server_wait(self.h, jid, float(timeout), ctypes.byref(info)))
        return info

    def mesh(self, jid, info):
        s = info.sizes
        rec = np.zeros((s.n_verts, 11), dtype=np.uint64)
        xyz = np.zeros((s.n_verts, 3), dtype=np.float64)
        self.check(serve_lib().rin_server_job_verts(self.h, jid, _p(rec, _u64p), _p(xyz, _f64p)))
        vs = np.zeros(s.n_faces + 1, dtype=np.uint64)
        v = np.zeros(s.n_face_verts, dtype=np.uint64)
        ts = np.zeros(s.n_faces + 1, dtype=np.uint64)
        tf = np.zeros((s.n_tet_faces, 2), dtype=np.uint64)
        fn = np.zeros((s.n_faces, 2), dtype=np.uint64)
        self.check(serve_lib().rin_server_job_faces(self.h, jid, _p(vs, _u64p), _p(v, _u64p), _p(ts, _u64p),
                                                    _p(tf, _u64p), _p(fn, _u64p)))
        return rec, xyz, vs, v, ts, tf, fn

    def release(self, jid):
        self.check(serve_lib().rin_server_job_release(self.h, jid))

    def stats(self):
        st = ServerStats()
        self.check(serve_lib().rin_server_stats_get(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if self.h:
            serve_lib().rin_server_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
