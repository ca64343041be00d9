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


@pytest.mark.skip(reason="concurrent slots are switched off: four queues running the thread-per-tet cut kernels "
                         "(up to 12.5 KB of private memory a lane) exhausted the runtime's scratch allocation "
                         "(HSA_STATUS_ERROR_OUT_OF_RESOURCES, even for a 3-wave dispatch); re-enable once those "
                         "kernels keep their complexes in LDS or device memory")
def test_results_same_alone_and_in_the_mix():
    import serving

    jobs = serving.make_jobs(16, base=100)
    mix, infos, meshes = serving.run_stream(jobs, rate=400.0, slots=4, keep=True)
    alone, _, ref = serving.run_stream(jobs, rate=0.0, slots=1, keep=True)
    assert mix["failed"] == 0 and alone["failed"] == 0
    assert mix["max_in_flight"] > 1, mix  # jobs overlapped on the device
    assert alone["max_in_flight"] == 1
    for m, r in zip(meshes, ref):
        _same(m, r)
    assert mix["p99_ms"] >= mix["p50_ms"] > 0


def test_admission_waits_for_memory():
    import serving

    jobs = serving.make_jobs(6, base=200)
    # a budget below one job's estimate admits one job at a time
    st, infos, _ = serving.run_stream(jobs, rate=0.0, slots=1, hbm_budget=1)
    assert st["failed"] == 0 and st["max_in_flight"] == 1
    starts = sorted(i.t_start for i in infos)
    ends = sorted(i.t_end for i in infos)
    assert all(s >= e - 1e-6 for s, e in zip(starts[1:], ends[:-1]))


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
