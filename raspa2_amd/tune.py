"""Autotuner of the screening service: for each framework-size bucket, the
cell-list cell size and the chunk (insertions per pipeline pass) that give
the highest Widom insertion rate on this GPU, written to a JSON table that
the service loads (screening.load_tuning).

    python -m raspa2_amd.tune [--out raspa2_amd/tuning.json]
"""
import argparse
import json
import os
import time

import numpy as np

from . import rhg as R
from . import workloads as W
from .screening import TUNING_PATH

BUCKETS = (1000, 2000, 4000, 8000)
CELLS = (2.0, 2.5, 3.0, 3.5, 4.0)
CHUNKS = (1 << 16, 1 << 18)


def representatives(per_bucket=3, scan=256):
    """Up to `per_bucket` screening frameworks in each size bucket."""
    out = {b: [] for b in BUCKETS}
    for i in range(scan):
        j = W.c3_job(i)
        for lo, hi in zip((0,) + BUCKETS[:-1], BUCKETS):
            if lo < j.framework.n <= hi and len(out[hi]) < per_bucket:
                out[hi].append(j)
        if all(len(v) >= per_bucket for v in out.values()):
            break
    return out


def rate(job, cell, chunk, n):
    ctx = job.framework.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
    ctx.set_culling(True, cell)
    ctx.set_widom_options(chunk=chunk, graphs=True)
    types = [ctx.add_pseudo_atom(*s) for s in job.probe.sites]
    mol = R.molecule(types, job.probe.body, start=job.probe.start, k_fb=job.probe.k_fb,
                     k_or=job.probe.k_or)
    ctx.widom_run(mol, job.seed, 2 * n, first=0, n=min(n, chunk))  # warm-up, sizes buffers
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.widom_run(mol, job.seed, 2 * n, first=n, n=n)
    dt = time.perf_counter() - t0
    ctx.close()
    return n / dt


def tune(n=1 << 18, per_bucket=3, log=print):
    reps = representatives(per_bucket)
    rows = []
    for b in BUCKETS:
        jobs = reps[b]
        if not jobs:
            continue
        best = None
        for cell in CELLS:
            for chunk in CHUNKS:
                # time for the bucket's frameworks, one request each
                t = sum(n / rate(j, cell, chunk, n) for j in jobs)
                log(f"bucket <= {b} atoms: cell {cell} chunk {chunk}: {len(jobs) * n / t:.4g} ins/s")
                if best is None or t < best[0]:
                    best = (t, cell, chunk)
        rows.append({"max_atoms": b, "cell_size": best[1], "chunk": best[2],
                     "rate": len(jobs) * n / best[0],
                     "frameworks": [j.index for j in jobs]})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--insertions", type=int, default=1 << 18)
    args = ap.parse_args()
    rows = tune(args.insertions)
    doc = {"what": "cell-list cell size [A] and chunk per framework-size bucket, mixed "
                   "precision, Ewald 1e-6, measured Widom insertion rate",
           "command": "python -m raspa2_amd.tune", "insertions_per_measurement": args.insertions,
           "buckets": rows}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc))


if __name__ == "__main__":
    main()
