"""Autotuner of the screening service: for each framework-size bucket, the
cell-list cell size, the chunk (insertions per pipeline pass) and the launch
shape of the culled site kernels (sites per lane of the LJ-only image-list
tiles, waves per workgroup) that serve the bucket's screening requests
fastest on this GPU, written to a JSON table that the service loads
(screening.load_tuning).

    python -m raspa2_amd.tune [--out raspa2_amd/tuning.json]

A candidate is scored by serving the bucket's representative requests (the
first screening frameworks of that size, each with its own probe and trial
count, 10^4 to 10^6 insertions) on a fresh service with `slots` requests in
flight: the rate counts context set-up (cell and image lists) together with
the insertions, as a screening run does.  The search is a coordinate descent
per bucket: cell size first (at the default tiles and chunk), then tiles,
then the chunk, each step keeping the best so far; after one warm-up serve,
each candidate is served three times and its best rate kept.
"""
import argparse
import json

from . import screening as SC
from . import workloads as W
from .screening import TUNING_PATH

BUCKETS = (1000, 2000, 4000, 8000)
CELLS = (2.0, 2.5, 3.0, 3.5)
TILES = ((4, 2), (4, 4), (2, 2), (4, 1), (2, 4))  # (sites per lane, waves per workgroup)
CHUNKS = (1 << 16, 1 << 17, 1 << 18)


def representatives(per_bucket=24, scan=4096):
    """Up to `per_bucket` screening frameworks in each size bucket, in the
    order of the screening draw (the first 256 hold none above 4000 atoms;
    the draw reaches about 6600)."""
    out = {b: [] for b in BUCKETS}
    for i in range(scan):
        j = W.c3_job(i)
        for lo, hi in zip((0,) + BUCKETS[:-1], BUCKETS):
            if lo < j.framework.n <= hi and len(out[hi]) < per_bucket:
                out[hi].append(j)
        if all(len(v) >= per_bucket for v in out.values()):
            break
    return out


def served_rate(jobs, bucket, cell, chunk, tiles, slots, repeats=3):
    """Insertions per second serving one request per framework of `jobs`
    with this bucket's candidate options (the best of `repeats` runs)."""
    index = {j.index: j for j in jobs}
    reqs = SC.screening_requests(jobs)
    row = (bucket, cell, chunk, tiles[0], tiles[1])
    best = 0.0
    for _ in range(repeats):
        out = SC.run(index, reqs, slots=slots, tuning=[row], slice_=chunk)
        s = SC.summary(out, reqs)
        if s["failed"]:
            raise RuntimeError(f"{s['failed']} requests failed while tuning")
        best = max(best, s["rate"])
    return best


def tune(per_bucket=24, slots=4, log=print):
    reps = representatives(per_bucket)
    rows = []
    for b in BUCKETS:
        jobs = reps[b]
        if not jobs:
            continue

        def score(cell, chunk, tiles):
            r = served_rate(jobs, b, cell, chunk, tiles, slots)
            log(f"bucket <= {b} atoms: cell {cell} chunk {chunk} tiles {tiles}: {r:.4g} ins/s", flush=True)
            return r

        best = {"cell": 2.5, "chunk": 1 << 18, "tiles": (4, 2)}
        served_rate(jobs, b, best["cell"], best["chunk"], best["tiles"], slots, repeats=1)  # warm-up
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
    ap.add_argument("--per-bucket", type=int, default=24)
    ap.add_argument("--slots", type=int, default=4)
    args = ap.parse_args()
    rows = tune(args.per_bucket, args.slots)
    doc = {"what": "per framework-size bucket: cell-list cell size [A], chunk, sites per lane (LJ-only lists) and "
                   "waves per workgroup of the site kernels; mixed precision, Ewald 1e-6; rate = insertions/s serving "
                   "the bucket's screening requests (set-up included) with the given slots",
           "command": f"python -m raspa2_amd.tune --per-bucket {args.per_bucket} --slots {args.slots}",
           "slots": args.slots, "buckets": rows}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc))


if __name__ == "__main__":
    main()
