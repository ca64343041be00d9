"""Serving stream: Poisson arrivals of mixed implicit-arrangement and
material-interface jobs (grid resolution in {16, 32, 64}, 2..32 functions of
mixed type, 70% / 30%), submitted to a rin.Server as they arrive; reports
latency percentiles beside throughput."""
import time

import numpy as np

import rin
import workloads as W


def make_jobs(n, base=0):
    jobs = []
    for i in range(n):
        res, fs, mi = W.serving_job(base + i)
        types, params = fs.arrays()
        jobs.append((mi, res, types, params))
    return jobs


def run_stream(jobs, rate, slots=4, seed=5, hbm_budget=0, keep=False):
    """Submit jobs at Poisson arrival times (rate jobs/s) and wait for all.
    Returns (summary dict, per-job infos, meshes if keep)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(jobs)) if rate > 0 else np.zeros(len(jobs))
    with rin.Server(0, slots=slots, hbm_budget=hbm_budget) as srv:
        t0 = srv.now()
        due = t0 + np.cumsum(gaps)
        ids = []
        for (mi, res, types, params), t in zip(jobs, due):
            delay = t - srv.now()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(mi, res, types, params))
        infos, meshes = [], []
        for jid in ids:
            info = srv.wait(jid)
            infos.append(info)
            if keep:
                meshes.append(srv.mesh(jid, info))
            srv.release(jid)
        st = srv.stats()
    lat = np.array([i.t_end - i.t_submit for i in infos])
    span = st.last_end - st.first_submit
    summary = {
        "jobs": len(jobs), "rate_jobs_per_s": rate, "slots": slots,
        "completed": int(st.completed), "failed": int(st.failed), "max_in_flight": int(st.max_in_flight),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "max_ms": float(lat.max() * 1e3), "mean_service_ms": float(np.mean([i.t_end - i.t_start for i in infos]) * 1e3),
        "jobs_per_s": len(jobs) / span if span > 0 else None,
        "tets_per_s": float(st.tets) / span if span > 0 else None,
        "hbm_budget_bytes": int(st.hbm_budget),
    }
    return summary, infos, meshes
