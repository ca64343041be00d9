"""Serving loop (include/rin/serve.h): each job's mesh is the same, bit for
bit, alone and in the mix of concurrently running jobs; admission by HBM
budget; a served job against the oracle."""
import numpy as np
import pytest

import support as S

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.view(np.uint64) if x.dtype == np.float64 else x,
                                      y.view(np.uint64) if y.dtype == np.float64 else y)


def test_results_same_alone_and_in_the_mix():
    """Jobs submitted faster than they finish run on four slots at once (their
    work overlaps on the device); every mesh equals, bit for bit, the same
    job run alone on one slot."""
    import serving

    jobs = serving.make_jobs(16, base=100)
    mix, infos, meshes = serving.run_stream(jobs, rate=400.0, slots=4, keep=True)
    alone, _, ref = serving.run_stream(jobs, rate=0.0, slots=1, keep=True)
    assert mix["failed"] == 0 and alone["failed"] == 0
    assert mix["max_in_flight"] > 1, mix  # jobs overlapped on the device
    assert alone["max_in_flight"] == 1
    # some pair of jobs ran at the same time on different slots
    assert any(a.t_start < b.t_end and b.t_start < a.t_end and a.slot != b.slot
               for i, a in enumerate(infos) for b in infos[i + 1:])
    for m, r in zip(meshes, ref):
        _same(m, r)
    assert mix["p99_ms"] >= mix["p50_ms"] > 0


def test_admission_by_memory_budget():
    """Four slots: with a budget below one job's estimate the server admits a
    job only when nothing else runs (start after the previous end); with the
    default budget the same stream overlaps."""
    import serving

    jobs = serving.make_jobs(8, base=200)
    st, infos, _ = serving.run_stream(jobs, rate=0.0, slots=4, hbm_budget=1)
    assert st["failed"] == 0 and st["max_in_flight"] == 1
    spans = sorted((i.t_start, i.t_end) for i in infos)
    assert all(b[0] >= a[1] - 1e-6 for a, b in zip(spans, spans[1:]))
    st2, _, _ = serving.run_stream(jobs, rate=0.0, slots=4)
    assert st2["failed"] == 0 and st2["max_in_flight"] > 1


@pytest.mark.parametrize("mi", [False, True])
def test_served_job_matches_oracle(mi):
    import rin
    import workloads as W

    fs = W.FuncSet()
    rng = np.random.default_rng(77 + mi)
    if mi:
        for _ in range(5):
            fs.sphere(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.0, 0.2))
    else:
        fs = W.mixed(rng, 2, 2, 1, 1)
    types, params = fs.arrays()
    with rin.Server(0, slots=1) as srv:
        jid = srv.submit(mi, 16, types, params)
        info = srv.wait(jid)
        rec, xyz, vs, v, ts, tf, fn = srv.mesh(jid, info)
    pts, tets = S.grid(16)
    ref = S.OracleRun(1 if mi else 0, pts, tets, S.evaluate(types, params, pts))
    np.testing.assert_array_equal(rec, ref.verts)
    np.testing.assert_array_equal(xyz.view(np.uint64), ref.xyz.view(np.uint64))
    np.testing.assert_array_equal(vs, ref.face_start)
    np.testing.assert_array_equal(v, ref.face_verts)
    np.testing.assert_array_equal(tf, ref.tf)
    np.testing.assert_array_equal(fn, ref.face_func)
