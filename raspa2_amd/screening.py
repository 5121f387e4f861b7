"""Screening workloads on the service (rhg_server_*): a catalogue of
synthetic frameworks, batched screening of all of them, and a stream of
Poisson-arriving requests.  The unit of work is RASPA2's Widom move
(src/mc_moves.c:7844); the estimators per request follow
src/statistics.c (rhg.widom_estimators)."""
import json
import os
import time
from dataclasses import dataclass

import numpy as np

from . import rhg as R
from . import serve as S
from . import workloads as W

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def load_tuning(path=TUNING_PATH):
    """[(max_atoms, cell_size, chunk, sites_per_lane, waves_per_block)] rows
    of the tuning table, or []."""
    if not os.path.exists(path):
        return []
    with open(path) as f:
        d = json.load(f)
    return [(int(r["max_atoms"]), float(r["cell_size"]), int(r["chunk"]), int(r.get("sites_per_lane", 4)),
             int(r.get("waves_per_block", 4))) for r in d["buckets"]]


def catalogue(n):
    """The n screening frameworks (with their default probe and trial count)."""
    return [W.c3_job(i) for i in range(n)]


@dataclass
class Request:
    framework: int
    probe: W.Probe
    n_trials: int
    seed: int
    arrival_s: float = 0.0


def screening_requests(jobs):
    """One request per framework, all present at the start."""
    return [Request(j.index, j.probe, j.n_trials, j.seed, 0.0) for j in jobs]


def poisson_requests(n_req, n_fw, rate, seed=4242):
    """n_req requests at Poisson arrivals of `rate` per second: a uniform
    framework, methane / CO2 / N2-like with equal probability, 10^4..10^6
    insertions log-uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    probes = (W.METHANE, W.CO2, W.N2LIKE)
    out = []
    for r in range(n_req):
        fw = int(rng.integers(n_fw))
        pr = probes[int(rng.integers(3))]
        n = int(round(10 ** rng.uniform(4.0, 6.0)))
        out.append(Request(fw, pr, n, 30_000 + r, float(t[r])))
    return out


def run(jobs, requests, slots=4, tuning=None, precision=R.PREC_MIXED, slice_=1 << 18,
        hbm_budget=0, live=False):
    """Serves `requests` over the frameworks of `jobs`; returns wall time,
    per-request results and the server counters.  live: the request loop
    starts first and a second thread submits each request at its arrival
    time (otherwise the requests are submitted beforehand with their arrival
    times)."""
    srv = S.Server(slots=slots, precision=precision, slice_=slice_, hbm_budget=hbm_budget)
    if tuning:
        srv.set_tuning(tuning)
    used = sorted({r.framework for r in requests})
    for k in used:
        srv.add_framework(k, jobs[k].framework)
    if live:
        import threading
        order = sorted(range(len(requests)), key=lambda k: requests[k].arrival_s)
        ids = [None] * len(requests)
        srv.start()
        t0 = time.perf_counter()

        def feed():
            for k in order:
                r = requests[k]
                delay = r.arrival_s - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(r.framework, r.probe, r.seed, r.n_trials)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        srv.stop()
        wall = time.perf_counter() - t0
    else:
        ids = [srv.submit(r.framework, r.probe, r.seed, r.n_trials, arrival_s=r.arrival_s)
               for r in requests]
        t0 = time.perf_counter()
        srv.run()
        wall = time.perf_counter() - t0
    res = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.close()
    return dict(wall=wall, results=res, stats=st)


def summary(out, requests):
    lat = np.array([r["latency_s"] for r in out["results"]])
    ok = np.array([r["status"] == 0 for r in out["results"]])
    total = sum(q.n_trials for q, good in zip(requests, ok) if good)
    return dict(insertions=int(total), wall_s=out["wall"], rate=total / out["wall"],
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                failed=int((~ok).sum()), requests=len(requests))
