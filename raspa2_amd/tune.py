"""Autotuner of the screening service: for each framework-size bucket, the
cell-list cell size, the chunk (insertions per pipeline pass) and the launch
shape of the culled site kernels (sites per lane of the LJ-only image-list
tiles, waves per workgroup) that give the highest Widom insertion rate on this GPU,
written to a JSON table that the service loads (screening.load_tuning).

    python -m raspa2_amd.tune [--out raspa2_amd/tuning.json]

The search is a coordinate descent per bucket: cell size first (at the
default tiles and chunk), then tiles, then the chunk, each step keeping the
best so far; every candidate is timed on the bucket's representative
frameworks (one request each, methane or CO2 as the screening draws them).
"""
import argparse
import json
import time

from . import rhg as R
from . import workloads as W
from .screening import TUNING_PATH

BUCKETS = (1000, 2000, 4000, 8000)
CELLS = (2.0, 2.5, 3.0, 3.5)
TILES = ((4, 2), (4, 4), (2, 2), (4, 1), (2, 4))  # (sites per lane, waves per workgroup)
CHUNKS = (1 << 16, 1 << 17, 1 << 18)


def representatives(per_bucket=3, scan=4096):
    """Up to `per_bucket` screening frameworks in each size bucket (the
    first 256 hold none above 4000 atoms; the draw reaches about 6600)."""
    out = {b: [] for b in BUCKETS}
    for i in range(scan):
        j = W.c3_job(i)
        for lo, hi in zip((0,) + BUCKETS[:-1], BUCKETS):
            if lo < j.framework.n <= hi and len(out[hi]) < per_bucket:
                out[hi].append(j)
        if all(len(v) >= per_bucket for v in out.values()):
            break
    return out


def rate(job, cell, chunk, tiles, n):
    ctx = job.framework.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
    ctx.set_culling(True, cell)
    ctx.set_widom_options(chunk=chunk, graphs=True)
    ctx.set_tiles(*tiles)
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

        def score(cell, chunk, tiles):
            # time for the bucket's frameworks, one request each
            t = sum(n / rate(j, cell, chunk, tiles, n) for j in jobs)
            r = len(jobs) * n / t
            log(f"bucket <= {b} atoms: cell {cell} chunk {chunk} tiles {tiles}: {r:.4g} ins/s")
            return r

        best = {"cell": 2.5, "chunk": 1 << 18, "tiles": (4, 2)}
        best_rate = score(best["cell"], best["chunk"], best["tiles"])
        for key, values in (("cell", CELLS), ("tiles", TILES), ("chunk", CHUNKS)):
            for v in values:
                if v == best[key]:
                    continue
                trial = dict(best, **{key: v})
                r = score(trial["cell"], trial["chunk"], trial["tiles"])
                if r > best_rate:
                    best, best_rate = trial, r
        rows.append({"max_atoms": b, "cell_size": best["cell"], "chunk": best["chunk"],
                     "sites_per_lane": best["tiles"][0], "waves_per_block": best["tiles"][1],
                     "rate": best_rate, "frameworks": [j.index for j in jobs],
                     "atoms": [j.framework.n for j in jobs]})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--insertions", type=int, default=1 << 18)
    args = ap.parse_args()
    rows = tune(args.insertions)
    doc = {"what": "per framework-size bucket: cell-list cell size [A], chunk, sites per lane (LJ-only lists) and waves per "
                   "workgroup of the site kernels; mixed precision, Ewald 1e-6, measured Widom insertion rate",
           "command": "python -m raspa2_amd.tune", "insertions_per_measurement": args.insertions,
           "buckets": rows}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc))


if __name__ == "__main__":
    main()
