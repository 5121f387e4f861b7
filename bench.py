#!/usr/bin/env python3
"""Widom insertion throughput of rigid CO2 in a 4608-atom charged framework.

Workload (one step): 10^6 Widom insertions (WidomAdsorbateMove,
src/mc_moves.c:7844) of TraPPE-like rigid CO2 into the 2x2x2 supercell
(51.6 A cubic, 4608 atoms, charges N(0,0.4 e) neutralised) with Lennard-Jones
(12 A, shifted) and Ewald (relative precision 1e-6, alpha 0.2651/A,
kmax 14, 6986 k-vectors); one first-bead trial and 10 orientation trials per
insertion, LJ biasing, T = 300 K.  Inputs are resident in HBM before the timed
region; trial positions come from the per-insertion counter stream on the
device.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload c1|c3|c4]

Prints one JSON line.  The CPU baseline is the oracle (the FP64 C
restatement) on every host core granted to the process over a bounded
sample of the same insertions.  Beside the C1 headline the line carries
`screening` (C3 and C4 below: served insertions/s, p50/p99 latency, the
unbatched plain GPU rate, an end-to-end roofline and the CPU baseline) and
`grid` (C2: the 0.1 A energy grid, nodes/s, roofline from the executed
pair evaluations, CPU baseline), then finite loading and flexible chains;
--headline-only leaves those out.

--workload c3: batched screening of 256 synthetic frameworks (637-3954
atoms, methane or CO2, 10^4-10^6 insertions each) through the screening
service; value = insertions/s over the whole screen, with p50/p99 of the
per-framework completion time.  --workload c4: Poisson arrivals of
requests (methane, CO2 or N2-like, any of the 256 frameworks) at 0.7x the
measured screening capacity, submitted from a second thread while the
request loop runs; p50/p99 request latency, and the same requests run one
at a time for comparison.  --workload c2: the grid summary with the
Ewald-Coulomb grid and the reference's grid decks.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Widom trial insertions/sec at stated rel. error; screening p50/p99 latency"
PEAK = {"mixed": 157.3, "fp64": 78.6}  # MI355X vector FP32 / FP64, TFLOP/s
FLOP_LJ_PAIR = 29      # minimum image + r^2 + cutoff test (21) + LJ (8)
FLOP_CC_EXTRA = 9      # real-space erfc Coulomb term on top of a pair
FLOP_CC_PAIR = 30      # Coulomb-only pair: 21 + 9
SEED = 20241203
FLOP_GRID_PAIR = 69    # grid node per (node, atom) pair: pair 29 + the derivative channels 40


def host_cores():
    """Host cores this process may run on (all of them serve the CPU baseline)."""
    return max(1, len(os.sched_getaffinity(0)))


def _cpu_worker(args):
    fw_arrays, start, count = args
    import numpy as np
    import oracle as O
    box, bc, x, y, z, q, eps, sig, mass, sites, body = fw_arrays
    s = O.System(box, bc, x, y, z, q, eps, sig, mass=mass, charge_method=O.CHARGE_EWALD,
                 ewald_precision=1e-6, temperature=300.0)
    types = [s.add_pseudo_atom(*site) for site in sites]
    mol = O.molecule(types, body, start=1, k_fb=1, k_or=10)
    t0 = time.perf_counter()
    rec = O.widom_run_stream(s, mol, SEED, start, count)
    dt = time.perf_counter() - t0
    return dt, rec


def cpu_baseline(fw, probe, n_per_worker):
    """Oracle (FP64 C restatement, RASPA2's all-pairs order and cost model)
    on the host cores: one core alone, then one process per core."""
    import multiprocessing as mp
    import numpy as np
    arrays = fw.oracle_args() + (probe.sites, probe.body)
    dt1, rec1 = _cpu_worker((arrays, 0, n_per_worker))
    single = n_per_worker / dt1
    cores = host_cores()
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(arrays, k * n_per_worker, n_per_worker) for k in range(cores)])
    wall = time.perf_counter() - t0
    total = cores * n_per_worker
    return dict(single=single, aggregate=total / wall, cores=cores, sample=total,
                rec0=rec1)


def _cpu_job_worker(args):
    job, count = args
    import oracle as O
    fw = job.framework
    s = O.System(fw.box, fw.bc, fw.xyz[:, 0], fw.xyz[:, 1], fw.xyz[:, 2], fw.charge, fw.eps_k,
                 fw.sigma, mass=fw.mass, charge_method=O.CHARGE_EWALD, ewald_precision=1e-6,
                 temperature=300.0)
    types = [s.add_pseudo_atom(*site) for site in job.probe.sites]
    mol = O.molecule(types, job.probe.body, start=job.probe.start, k_fb=job.probe.k_fb,
                     k_or=job.probe.k_or)
    t0 = time.perf_counter()
    O.widom_run_stream(s, mol, job.seed, 0, count)
    return time.perf_counter() - t0


def cpu_baseline_screening(jobs, count):
    """Oracle on the host cores: `count` insertions of each of `jobs`, one
    process per core; rate = insertions / summed process time x cores."""
    import multiprocessing as mp
    cores = host_cores()
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        times = pool.map(_cpu_job_worker, [(j, count) for j in jobs])
    wall = time.perf_counter() - t0
    total = count * len(jobs)
    return dict(aggregate=total / wall, single=total / sum(times), cores=cores, sample=total)


def _grid_worker(args):
    arrays, geom, unit_box, inv, pts = args
    import oracle as O
    box, bc, x, y, z, q, eps, sig, mass = arrays
    s = O.System(box, bc, x, y, z, q, eps, sig, mass=mass)
    t = s.add_pseudo_atom(148.0, 3.73, 0.0)
    t0 = time.perf_counter()
    out = [s.interpolate_direct(t, geom, O.BC_TRICLINIC, unit_box, inv, p) for p in pts]
    return out, time.perf_counter() - t0


def grid_cpu_reference(n_points, cores=None):
    """The oracle's FP64 InterpolateVDWGrid over corner nodes computed on the
    fly at n_points seeded off-grid points of the C2 cell (0.1 A grid), one
    process per host core; run before this process touches the GPU (forked
    workers).  Returns the points, values and the node rate."""
    import multiprocessing as mp
    import numpy as np
    import oracle as O
    from raspa2_amd import workloads as W
    fw = W.c2_framework()
    geom = O.grid_geometry(fw.unit_box, 0.1)
    inv, _ = O.invert3x3(fw.unit_box)
    rng = np.random.Generator(np.random.PCG64(SEED))
    pts = rng.uniform(0.0, 1.0, (n_points, 3)) @ np.asarray(fw.unit_box).reshape(3, 3)
    cores = cores or host_cores()
    parts = np.array_split(np.arange(n_points), cores)
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_grid_worker, [(fw.oracle_args(), geom, fw.unit_box, inv, pts[ix]) for ix in parts])
    wall = time.perf_counter() - t0
    return {"points": pts, "u": np.concatenate([np.asarray(r[0]) for r in res]),
            "cpu_baseline": {"value": round(8 * n_points / wall, 1), "unit": "grid nodes/s", "cores": cores,
                             "kind": "port",
                             "sample": f"{8 * n_points} nodes (the 8 corners of {n_points} seeded off-grid "
                                       "points), oracle FP64 all-pairs, one process per core"}}


def grid_measure(ref):
    """C2 on the device: the 0.1 A energy grid of a methane-like probe over
    the triclinic 26 A / 80 deg cell (1200 atoms), plain FP64 and mixed
    builds (FP32 node storage), each checked against the oracle's values at
    the reference points; the mixed build's roofline from the (node, atom or
    image) evaluations it executed."""
    import numpy as np
    import oracle as O
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W
    fw = W.c2_framework()
    pts, u_ref = ref["points"], ref["u"]
    e2k = O.units()["energy_to_kelvin"]
    out = {}
    for name, prec in (("fp64", R.PREC_FP64), ("mixed", R.PREC_MIXED)):
        ctx = fw.context(precision=prec)
        t = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
        ctx.MakeGrid(t, fw.unit_box, 0.5).close()  # warm-up: cell and image lists, code objects
        ctx.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        g = ctx.MakeGrid(t, fw.unit_box, 0.1)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        _, ms = ctx.profile_get("grid_build")
        ctx.profile(False)
        nodes = int(np.prod(np.asarray(g.geom[:3]) + 1))
        u = g.interpolate(pts)
        capped_same = bool(np.array_equal(u >= 1e7, u_ref >= 1e7))
        ok = u_ref < 1e7
        d = np.abs(u[ok] - u_ref[ok]) * e2k
        rel = d / np.maximum(np.abs(u_ref[ok]) * e2k, 1.0)
        peak = PEAK["mixed" if prec == R.PREC_MIXED else "fp64"]
        ach = g.exec_pairs * FLOP_GRID_PAIR / (ms * 1e-3) / 1e12 if ms else 0.0
        out[name] = {"nodes": nodes, "nodes_per_s": round(nodes / wall, 1), "wall_s": round(wall, 5),
                     "kernel_ms": round(ms, 4), "executed_pairs": g.exec_pairs,
                     "max_rel_err": float(rel.max()), "max_abs_err_K": float(d.max()),
                     "capped_flags_identical": capped_same, "samples_checked": int(ok.sum()),
                     "roofline": {"bound": "valu", "kernel": "grid_build", "achieved": round(ach, 3),
                                  "peak": peak, "unit": "TFLOP/s", "frac": round(ach / peak, 4),
                                  "note": f"{FLOP_GRID_PAIR} flops per executed (node, atom) evaluation "
                                          "over the HIP-event time of the build launches"}}
        g.close()
        ctx.close()
    return {"workload": "C2 energy grid: methane-like LJ probe, triclinic 26 A 80 deg cell, 1200 atoms, "
                        "0.1 A spacing (352 x 294 x 254 nodes), 8 channels, FP32 storage",
            "value": out["mixed"]["nodes_per_s"], "unit": "grid nodes/s",
            "plain_fp64": out["fp64"], "mixed": out["mixed"], "roofline": out["mixed"]["roofline"],
            "check": f"{len(pts)} seeded off-grid points: device interpolation vs the oracle's "
                     "InterpolateVDWGrid over corner nodes computed on the fly (FP64)",
            "cpu_baseline": ref["cpu_baseline"]}


def grid_bench(args):
    """--workload c2: the grid summary plus the Ewald-Coulomb grid and the
    reference's grid decks."""
    ref = grid_cpu_reference(args.grid_samples if not args.no_cpu else 2000)
    g = grid_measure(ref)
    line = {"metric": "energy-grid nodes/s (MakeGrid with derivatives)", "value": g["value"],
            "unit": "grid nodes/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": round(g["mixed"]["wall_s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 pair terms, fp64 accumulation", "data": "synthetic",
            "config": {"workload": g["workload"], "nodes": g["mixed"]["nodes"], "spacing_A": 0.1,
                       "parallelism": "single GPU"},
            "plain_fp64": g["plain_fp64"], "mixed": g["mixed"], "roofline": g["roofline"],
            "coulomb_grid": coulomb_grid_bench(args), "grid_decks": grid_deck_bench(args),
            "check": g["check"], "cpu_baseline": g["cpu_baseline"]}
    print(json.dumps(line))


def coulomb_grid_bench(args):
    """Ewald-Coulomb grid of the charged 25.8 A cubic cell (2x2x2 supercell,
    4608 atoms) at 0.1 A: FP64 build, nodes checked against the oracle's at
    seeded node indices, and the RASPA .grid file written and read back."""
    import tempfile
    import numpy as np
    import oracle as O
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W
    fw = W.c1_framework()
    spacing = 0.1
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    g = ctx.MakeCoulombGrid(fw.unit_box, 0.5)  # warm-up: cell lists, code objects
    g.close()
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    g = ctx.MakeCoulombGrid(fw.unit_box, spacing)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    _, ms = ctx.profile_get("grid_build")
    ctx.profile(False)
    nodes = int(np.prod(np.asarray(g.geom[:3]) + 1))
    data = g.download()
    geom = O.grid_geometry(fw.unit_box, spacing)
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    sys_o = O.System(box, bc, x, y, z, q, eps, sig, mass=mass, charge_method=O.CHARGE_EWALD)
    rng = np.random.Generator(np.random.PCG64(SEED + 2))
    idx = [tuple(int(rng.integers(0, s)) for s in data.shape[:3]) for _ in range(64)]
    # CPU baseline leg: the oracle's FP64 nodes at the sampled indices, timed
    # on one core; the same values check the device grid
    t0 = time.perf_counter()
    refs = [sys_o.coulomb_grid_node(geom, i, j, k) for (i, j, k) in idx]
    cpu_s = time.perf_counter() - t0
    worst = 0.0
    capped_same = True
    for (i, j, k), ref in zip(idx, refs):
        got = data[i, j, k]
        capped_same &= bool((ref[0] == np.float32(2e7)) == (got[0] == np.float32(2e7)))
        ulp = np.spacing(np.abs(ref)).astype(np.float64)
        worst = max(worst, float(np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(ulp, 1e-30))))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c1_Electrostatics_Ewald.grid")
        t0 = time.perf_counter()
        g.write(path)
        t_w = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        h = ctx.ReadGrid(path, fw.unit_box, coulomb=True, n_unit_cells=(2, 2, 2))
        t_r = time.perf_counter() - t0
        same = bool(np.array_equal(h.download().view(np.uint32), data.view(np.uint32)))
        h.close()
    g.close()
    ctx.close()
    return {"workload": "Ewald-Coulomb grid (unit charge, real space, alpha from 1e-6), 25.8 A cubic "
                        "cell of the 2x2x2 supercell, 4608 atoms, 0.1 A spacing, FP64 compute, FP32 storage",
            "nodes": nodes, "nodes_per_s": nodes / wall, "wall_s": wall, "kernel_ms": ms,
            "check": f"{len(idx)} seeded nodes vs the oracle: cap flags equal {capped_same}, "
                     f"worst channel difference {worst:.1f} float32 ulps",
            "file": {"bytes": size, "write_s": t_w, "read_s": t_r, "read_back_bit_identical": same},
            "cpu_baseline": {"value": round(len(idx) / cpu_s, 1), "unit": "grid nodes/s", "cores": 1,
                             "kind": "port", "sample": f"{len(idx)} seeded nodes, oracle FP64 all-pairs"}}


def grid_deck_bench(args):
    """The reference's MakingGrids/UsingGrids example decks: IRMOF-1 with the
    Dubbeldam2007FlexibleIRMOF-1 force field, C_co2 and O_co2 grids and the
    Ewald-Coulomb grid at 0.1 A, written under RASPA2's file names and read
    back; then CO2 Widom insertion served from the grids against the direct
    (mixed) sums over the same trials."""
    import tempfile
    import numpy as np
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W
    fw, probe, D = W.irmof1_dubbeldam(os.path.join(ROOT, "tests", "golden"))
    sp = D["spacing_vdw"]

    def context(prec):
        c = fw.context(charge_method=R.CHARGE_EWALD, ewald_precision=D["ewald_precision"],
                       temperature=D["temperature"], precision=prec)
        types = [c.add_pseudo_atom(*s) for s in probe.sites]
        return c, types, R.molecule(types, probe.body, start=probe.start, k_fb=probe.k_fb,
                                    k_or=probe.k_or)

    names = [n for n, _ in D["co2"]]
    ctx, types, mol = context(R.PREC_FP64)
    for g in (ctx.MakeGrid(types[0], fw.unit_box, 0.5), ctx.MakeCoulombGrid(fw.unit_box, 0.5)):
        g.close()  # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    grids = {n: ctx.MakeGrid(types[names.index(n)], fw.unit_box, sp) for n in D["grid_types"]}
    coul = ctx.MakeCoulombGrid(fw.unit_box, D["spacing_coulomb"])
    ctx.synchronize()
    t_build = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as root:
        t0 = time.perf_counter()
        paths = {}
        for n, g in grids.items():
            paths[n] = R.grid_path(root, D["forcefield"], D["framework"], sp, pseudo_atom=n,
                                   shifted=D["shifted"])
            os.makedirs(os.path.dirname(paths[n]), exist_ok=True)
            g.write(paths[n])
        pc = R.grid_path(root, D["forcefield"], D["framework"], D["spacing_coulomb"],
                         n_unit_cells=D["unit_cells"])
        os.makedirs(os.path.dirname(pc), exist_ok=True)
        coul.write(pc)
        t_write = time.perf_counter() - t0
        for g in list(grids.values()) + [coul]:
            g.close()
        t0 = time.perf_counter()
        read = {n: ctx.ReadGrid(p, fw.unit_box) for n, p in paths.items()}
        rc = ctx.ReadGrid(pc, fw.unit_box, coulomb=True, n_unit_cells=D["unit_cells"])
        t_read = time.perf_counter() - t0
    for j, n in enumerate(names):
        ctx.use_grid(types[j], read[n])
    ctx.use_coulomb_grid(rc)
    direct, dtypes, dmol = context(R.PREC_MIXED)
    n = 1 << 20
    rates = {}
    for name, c, m in (("grids", ctx, mol), ("direct_mixed", direct, dmol)):
        c.widom_run(m, SEED, 4 * n, first=0, n=n)
        c.synchronize()
        t0 = time.perf_counter()
        for s in range(3):
            c.widom_run(m, SEED, 4 * n, first=(s + 1) * n, n=n)
        c.synchronize()
        rates[name] = 3 * n / (time.perf_counter() - t0)
    _, rec = ctx.widom_run(mol, SEED, 200000, records=True)
    _, rec0 = direct.widom_run(dmol, SEED, 200000, records=True)
    ratio = float(rec[:, 0].mean() / rec0[:, 0].mean())
    nodes = int(np.prod(np.asarray(rc.geom[:3]) + 1))
    for g in list(read.values()) + [rc]:
        g.close()
    ctx.close()
    direct.close()
    return {"workload": "examples/Auxiliary/MakingGrids + UsingGrids: IRMOF-1 (424 atoms, 1x1x1), "
                        "Dubbeldam2007FlexibleIRMOF-1, grids C_co2, O_co2 and Ewald-Coulomb at 0.1 A; "
                        "TraPPE CO2 Widom (k_fb=1, k_or=10) at 298 K",
            "nodes_per_grid": nodes,
            "build_3_grids_s": t_build, "write_s": t_write, "read_s": t_read,
            "widom_insertions_per_s": {k: round(v, 1) for k, v in rates.items()},
            "w_ratio_grids_over_direct": ratio,
            "w_check": "mean Widom weight over the same 200000 insertions, grids vs direct mixed sums"}


def screening_cpu(args, jobs):
    """CPU baseline of the screening workloads (before the GPU is touched):
    the oracle on the first screen_cpu_sample insertions of each of the
    first 16 screening requests, one process per host core."""
    if args.no_cpu:
        return None
    count = args.screen_cpu_sample
    cpu = cpu_baseline_screening(jobs[:min(16, len(jobs))], count)
    return {"value": round(cpu["aggregate"], 2), "unit": "insertions/s", "cores": cpu["cores"],
            "kind": "port", "single_core": round(cpu["single"], 2),
            "sample": f"{cpu['sample']} insertions: the first {count} of each of the first 16 screening "
                      "requests (their own frameworks and probes), oracle FP64 C restatement, one process "
                      "per core"}


def _served_roofline(stats, wall):
    """Executed-pair flops of the culled site kernels over the served wall
    time (every kernel, set-up and host work included)."""
    flops = (stats["exec_lj_pairs"] * FLOP_LJ_PAIR + stats["exec_ljcc_pairs"] * (FLOP_LJ_PAIR + FLOP_CC_EXTRA)
             + stats["exec_cc_pairs"] * FLOP_CC_PAIR)
    ach = flops / wall / 1e12
    return {"bound": "valu", "kernel": "site_energy (end to end)", "achieved": round(ach, 3), "peak": PEAK["mixed"],
            "unit": "TFLOP/s", "frac": round(ach / PEAK["mixed"], 4),
            "note": "algorithmic flops of the pair evaluations the culled site kernels executed for the "
                    "served requests, over the wall time of the whole served run (set-up, every kernel "
                    "and host work included), against the packed FP32 vector peak"}


def screening_measure(args, jobs, cpu_obj, which=("c3", "c4")):
    """C3 (one request per framework, all present at t=0) and C4 (Poisson
    arrivals at 0.7x the measured C3 request capacity, methane / CO2 /
    N2-like) through the screening service, each against the same requests
    run one at a time (unbatched plain GPU path)."""
    from raspa2_amd import screening as SC
    tuning = SC.load_tuning()
    reqs = SC.screening_requests(jobs)
    # bring the device up (not timed): a short screen of four frameworks
    warm = [SC.Request(r.framework, r.probe, 10_000, r.seed) for r in reqs[:4]]
    SC.run(jobs, warm, slots=args.slots, tuning=tuning)
    c3 = SC.run(jobs, reqs, slots=args.slots, tuning=tuning)
    s3 = SC.summary(c3, reqs)
    out = {}
    if "c3" in which:
        serial = SC.run(jobs, reqs, slots=1, tuning=tuning)
        s1 = SC.summary(serial, reqs)
        out["c3"] = {
            "workload": f"C3 batched screening: {len(jobs)} synthetic frameworks (637-3954 atoms, "
                        "charged, Ewald 1e-6), methane or CO2, 10^4-10^6 insertions each, mixed precision",
            "value": round(s3["rate"], 1), "unit": "insertions/s", "wall_s": round(s3["wall_s"], 4),
            "frameworks": len(jobs), "insertions": s3["insertions"], "slots": args.slots,
            "latency": {"p50_s": s3["p50_s"], "p99_s": s3["p99_s"],
                        "what": "completion time of each framework's request, all submitted at t=0"},
            "unbatched_plain_gpu": {"value": round(s1["rate"], 1), "p50_s": s1["p50_s"], "p99_s": s1["p99_s"],
                                    "what": "same requests one at a time"},
            "roofline": _served_roofline(c3["stats"], s3["wall_s"]),
            "cpu_baseline": cpu_obj, "server": c3["stats"], "failed": s3["failed"], "tuning": tuning}
    if "c4" in which:
        capacity = len(reqs) / s3["wall_s"]
        rate = 0.7 * capacity
        r4 = SC.poisson_requests(args.requests, len(jobs), rate)
        c4 = SC.run(jobs, r4, slots=args.slots, tuning=tuning, live=True)
        s4 = SC.summary(c4, r4)
        plain_reqs = [SC.Request(r.framework, r.probe, r.n_trials, r.seed, 0.0) for r in r4]
        plain = SC.run(jobs, plain_reqs, slots=1, tuning=tuning)
        sp = SC.summary(plain, plain_reqs)
        out["c4"] = {
            "workload": f"C4 serving: {len(r4)} Poisson requests at {rate:.2f}/s (0.7 x measured screening "
                        f"capacity {capacity:.2f}/s) over {len(jobs)} frameworks, methane / CO2 / N2-like, "
                        "10^4-10^6 insertions each, mixed precision, submitted live from a second thread",
            "value": round(s4["rate"], 1), "unit": "insertions/s", "wall_s": round(s4["wall_s"], 4),
            "requests": len(r4), "arrival_rate_per_s": rate, "insertions": s4["insertions"],
            "slots": args.slots,
            "latency": {"p50_s": s4["p50_s"], "p99_s": s4["p99_s"], "what": "arrival to completion of each request"},
            "unbatched_plain_gpu": {"value": round(sp["rate"], 1),
                                    "what": "same requests one at a time, all present at the start"},
            "roofline": _served_roofline(c4["stats"], s4["wall_s"]),
            "cpu_baseline": cpu_obj, "server": c4["stats"], "failed": s4["failed"]}
    return out


def screening_bench(args):
    """--workload c3 / c4: one JSON line for the screening workload."""
    from raspa2_amd import screening as SC
    jobs = SC.catalogue(args.frameworks)
    cpu_obj = screening_cpu(args, jobs)
    r = screening_measure(args, jobs, cpu_obj, which=(args.workload,))[args.workload]
    out = {"metric": METRIC, "value": r["value"], "unit": "insertions/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(r["wall_s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32 pair terms, fp64 accumulation/Ewald", "data": "synthetic",
           "config": {"workload": r["workload"], "parallelism": "single GPU"},
           "latency": r["latency"], "unbatched": r["unbatched_plain_gpu"], "roofline": r["roofline"],
           "cpu_baseline": r["cpu_baseline"], "detail": r}
    print(json.dumps(out))


def loaded_bench(fw, probe, seed=SEED + 3, n_ads=100):
    """Finite loading on the C1 supercell: n_ads CO2 molecules at the
    lowest-energy of 10^5 random centres and orientations (host LJ energy
    below 1e6, far from the 1e7 overlap threshold; centres at least 4 A apart),
    then CO2 Widom insertion with the guest-guest terms and the Fourier change
    against the stored adsorbate sums, and the retrace (CBMC deletion weight)
    of every adsorbed molecule."""
    import numpy as np
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W
    ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
    types, mol = W.register_probe(ctx, probe)
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.asarray(fw.box).reshape(3, 3)
    cand = rng.uniform(0.0, 1.0, (100000, 3)) @ m
    v = rng.normal(size=(100000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    body = np.asarray(probe.body)
    sites = [cand + body[j][2] * v for j in range(len(body))]  # rigid along its axis
    u = np.zeros(len(cand))
    for j, t in enumerate(types):
        uv, _, _ = ctx.energy_at_positions(sites[j], t)
        u += uv
    L = np.diag(m)
    chosen = []
    for i in np.argsort(u):
        if u[i] >= 1e6 or len(chosen) == n_ads:
            break
        d = cand[chosen] - cand[i] if chosen else np.zeros((0, 3))
        d -= L * np.round(d / L)
        if len(chosen) == 0 or np.min(np.sum(d * d, axis=1)) >= 16.0:
            chosen.append(i)
    xyz = np.concatenate([np.stack([sites[j][i] for j in range(len(body))]) for i in chosen])
    start = list(range(0, 3 * len(chosen) + 1, 3))
    ctx.set_adsorbates(start, list(types) * len(chosen), xyz)
    n = 1 << 20
    rates = {}
    ctx.widom_run(mol, SEED, 4 * n, first=0, n=n)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(3):
        ctx.widom_run(mol, SEED, 4 * n, first=(s + 1) * n, n=n)
    ctx.synchronize()
    rates["widom_insertions_per_s"] = 3 * n / (time.perf_counter() - t0)
    blocks, _ = ctx.widom_run(mol, SEED, n, first=0, n=n)
    info = ctx.info()
    N = len(chosen)
    est = R.widom_estimators(blocks, 300.0, info["density"])
    # retrace every adsorbed molecule many times over (fresh random trials)
    reps = 2000
    olds = np.repeat(xyz.reshape(N, 3, 3), reps, axis=0)
    fb = rng.uniform(0.0, 1.0, (N * reps, probe.k_fb, 3)) @ m
    cord = np.zeros((N * reps, probe.k_or, 3, 3))
    for iu in range(probe.k_or):
        w = rng.normal(size=(N * reps, 3))
        w /= np.linalg.norm(w, axis=1)[:, None]
        for j in range(3):
            cord[:, iu, j] = (body[j][2] - body[probe.start][2]) * w
    excl = np.repeat(np.arange(N, dtype=np.int32), reps)
    ctx.RetraceMolecule(mol, excl[:1000], olds[:1000], fb[:1000], cord[:1000])  # warm-up
    t0 = time.perf_counter()
    rec = ctx.RetraceMolecule(mol, excl, olds, fb, cord)
    rates["retrace_molecules_per_s"] = len(olds) / (time.perf_counter() - t0)
    ctx.close()
    return {"workload": f"C1 supercell with {N} CO2 adsorbed (lowest-energy, non-overlapping "
                        "positions), mixed precision: CO2 Widom insertion with guest-guest terms "
                        "and the adsorbate Fourier sums; retraces of the adsorbed molecules "
                        f"({reps} random trial sets each, each molecule excluded from its own "
                        "guest-guest sums)",
            "n_adsorbates": N, **{k: round(v, 1) for k, v in rates.items()},
            "w_overall": est["w_overall"],
            "retrace_accepted_fraction": float(np.mean(rec[:, 13] == 4))}


def chain_bench(no_cpu=False, n=1 << 18, steps=3, cpu_sample=1500):
    """n-hexane (TraPPE united atom) Widom insertion by CBMC with the internal
    MC scheme in IRMOF-1 (the reference's CIF in P1 with the
    Dubbeldam2007FlexibleIRMOF-1 LJ parameters, 424 atoms, 25.83 A cubic,
    from the committed fixtures), LJ 12 A shifted, 300 K, mixed precision,
    k_fb = 10, k_trial = 10, k_torsion = 100, 150 moves per open bead.  The
    ideal-gas Rosenbluth weight comes from the same growth in the empty cell.
    The oracle times a bounded sample of the same insertions on one core, and
    the FP64 GPU path is compared with it insertion by insertion."""
    import numpy as np
    import oracle as O
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W
    fw, _, _ = W.irmof1_dubbeldam(os.path.join(ROOT, "tests", "golden"))
    probe = W.trappe_alkane(6)
    ctx = fw.context(precision=R.PREC_MIXED)
    _, ch = W.register_chain(ctx, probe)
    total = (steps + 1) * n
    ctx.widom_chain_run(ch, SEED, total, first=0, n=n)
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    blocks = R.WidomBlocks()
    t0 = time.perf_counter()
    for s in range(steps):
        ctx.widom_chain_run(ch, SEED, total, first=(s + 1) * n, n=n, blocks=blocks)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    kernels = {}
    for name in ("chain_grow", "chain_trials", "site_energy", "bin", "chain_select", "widom_gen",
                 "block_reduce"):
        cnt, ms = ctx.profile_get(name)
        kernels[name] = {"launches": cnt, "total_ms": round(ms, 4),
                         "avg_ms": round(ms / cnt, 5) if cnt else None}
    ctx.profile(False)
    e = np.zeros(0)
    empty = R.Context(fw.box, fw.bc, e, e, e, e, e, charge=e, mass=e, precision=R.PREC_FP64)
    _, che = W.register_chain(empty, probe)
    bi, _, _ = empty.widom_chain_run(che, SEED + 1, 1 << 18)
    b = bi.as_array()
    w_ig = b[:, 1].sum() / b[:, 0].sum()
    empty.close()
    est = R.widom_estimators(blocks, 300.0, ctx.info()["density"], ideal_gas_weight=w_ig)
    # plain FP64 GPU rate and per-insertion parity with the oracle on a sample
    ref = fw.context(precision=R.PREC_FP64)
    _, ch64 = W.register_chain(ref, probe)
    ref.widom_chain_run(ch64, SEED, total, first=n, n=n)
    ref.synchronize()
    t2 = time.perf_counter()
    ref.widom_chain_run(ch64, SEED, total, first=n, n=n)
    ref.synchronize()
    plain = n / (time.perf_counter() - t2)
    out = {"workload": "n-hexane (TraPPE UA, 6 beads) CBMC Widom insertion in IRMOF-1 (424 atoms, "
                       "25.83 A cubic), LJ 12 A shifted, 300 K, mixed precision, k_fb=10, k_trial=10, "
                       "k_torsion=100, 150 internal moves per open bead",
           "insertions_per_step": n, "steps": steps,
           "insertions_per_s": round(steps * n / elapsed, 1),
           "fp64_plain_gpu_insertions_per_s": round(plain, 1),
           "kernels": kernels,
           "ideal_gas_rosenbluth_weight": w_ig,
           "estimators": {"henry_mol_kg_pa": est["kh_overall"], "u_gh_K": est["u"],
                          "w": est["w_overall"]}}
    if not no_cpu:
        s_o = O.System(*fw.oracle_args()[:2], *fw.oracle_args()[2:8], mass=fw.mass, cutoff_vdw=12.0,
                       shifted=True, temperature=300.0)
        types = [s_o.add_pseudo_atom(ep, sg, q) for ep, sg, q in probe.sites]
        co = O.chain(types, probe.bonds, probe.bends, probe.torsions, probe.intra, start=probe.start,
                     k_fb=probe.k_fb, k_trial=probe.k_trial, k_torsion=probe.k_torsion, moves=probe.moves)
        t3 = time.perf_counter()
        rec_o = O.chain_run_stream(s_o, co, SEED, n, cpu_sample)
        cpu_rate = cpu_sample / (time.perf_counter() - t3)
        _, rec_g, _ = ref.widom_chain_run(ch64, SEED, total, first=n, n=cpu_sample, records=True)
        st_o, st_g = rec_o[:, O.CREC["STAGE"]], rec_g[:, R.CREC["STAGE"]]
        ok = (st_o == 4) & (st_g == 4)
        werr = float(np.max(np.abs(rec_g[ok, 0] - rec_o[ok, 0]) / rec_o[ok, 0])) if ok.any() else None
        out["cpu_baseline"] = {"value": round(cpu_rate, 2), "unit": "insertions/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{cpu_sample} insertions of the same workload (insertions "
                                         f"{n}..{n + cpu_sample - 1} of the job's counter stream), oracle "
                                         "FP64 C restatement, one core"}
        out["cpu_parity_fp64"] = {"insertions": cpu_sample, "stages_identical": bool(np.array_equal(st_o, st_g)),
                                  "max_rel_err_W": werr}
    ref.close()
    ctx.close()
    return out


FLOP_PSD_PAIR = 31  # per (sphere centre, atom): 3 sub, rectangular minimum image 18, r^2 5, cut test 2, sqrt, sub, min


def psd_bench(no_cpu=False):
    """The reference's pore-size distribution example (examples/Auxiliary/PSD/
    IRMOF-1): IRMOF-1 with GenericMOFs sigmas, PSDProbeDistance Sigma, 10^4
    cycles of 2 x 10^5 sphere centres each, 1000 bins over 10 A, FP64.  Rate
    in sphere centres per second (each tested against all 424 atoms); the
    oracle runs three cycles of the same job on one core."""
    import json as _json
    import numpy as np
    import oracle as O
    from raspa2_amd import frontend as F
    from raspa2_amd import rhg as R
    gold = os.path.join(ROOT, "tests", "golden")
    G = _json.load(open(os.path.join(gold, "psd_irmof1.json")))
    ff = F.ForceField.load(os.path.join(gold, "share", "raspa", "forcefield", "GenericMOFs"))
    cif = open(os.path.join(gold, "share", "raspa", "structures", "cif", "IRMOF-1.cif")).read()
    fw = F.framework_from_cif(cif, ff)
    ctx = fw.context(precision=R.PREC_FP64)
    n = G["cycles"]
    ctx.PoreSizeDistribution(200, 20 * n, seed=2, bins=G["bins"], range_=G["range"])  # warm-up
    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    hist = ctx.PoreSizeDistribution(n, 20 * n, seed=12, bins=G["bins"], range_=G["range"])
    ctx.synchronize()
    el = time.perf_counter() - t0
    cnt, ms = ctx.profile_get("psd")
    ctx.profile(False)
    centres = float(n) * 20 * n
    pairs = centres * fw.n
    ours, ref = hist / n, np.array(G["cumulative"])
    out = {"workload": "IRMOF-1 pore-size distribution (examples/Auxiliary/PSD/IRMOF-1): 424 atoms, "
                       "GenericMOFs sigma, probe distance Sigma, 10^4 cycles x 2x10^5 sphere centres, "
                       "1000 bins over 10 A, FP64",
           "seconds": round(el, 3), "centres_per_s": round(centres / el, 1),
           "centre_atom_pairs_per_s": round(pairs / el, 1),
           "roofline": {"bound": "valu", "kernel": "k_psd_b", "achieved": round(pairs * FLOP_PSD_PAIR / (ms * 1e-3) / 1e12, 3),
                        "peak": 78.6, "unit": "TFLOP/s (FP64)",
                        "frac": round(pairs * FLOP_PSD_PAIR / (ms * 1e-3) / 1e12 / 78.6, 4),
                        "note": f"{FLOP_PSD_PAIR} flops per centre-atom pair over the HIP-event time of the PSD launches"},
           "void_fraction": float(ours[0]), "void_fraction_printed": float(ref[0]),
           "max_abs_diff_cumulative": float(np.max(np.abs(ours - ref)))}
    if not no_cpu:
        s = O.System(fw.box, fw.bc, fw.xyz[:, 0], fw.xyz[:, 1], fw.xyz[:, 2], fw.charge, fw.eps_k, fw.sigma,
                     mass=fw.mass, cutoff_vdw=12.0)
        t1 = time.perf_counter()
        O.psd_histogram(s, 1.0, 20 * n, 12, 0, 3, G["bins"], G["range"])
        cpu = 3 * 20 * n / (time.perf_counter() - t1)
        out["cpu_baseline"] = {"value": round(cpu, 1), "unit": "sphere centres/s", "cores": 1, "kind": "port",
                               "sample": "cycles 0..2 of the same job (6x10^5 centres), oracle FP64 C restatement"}
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--insertions", type=int, default=1_000_000, help="insertions per step")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cpu-sample", type=int, default=2000, help="C1 insertions per CPU process")
    ap.add_argument("--screen-cpu-sample", type=int, default=2000,
                    help="insertions of each of 16 screening requests in their CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "chain", "psd"], default="c1")
    ap.add_argument("--grid-samples", type=int, default=100_000, help="c2 off-grid checks")
    ap.add_argument("--grid-check", type=int, default=10_000,
                    help="off-grid checks of the grid summary in the default line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 512 MiB device copy kernel (k_stream_copy) three "
                         "times, to scale profiler byte counters")
    ap.add_argument("--headline-only", action="store_true",
                    help="c1: leave out the screening, grid, finite-loading and flexible-chain "
                         "legs (profiles of the headline steps)")
    ap.add_argument("--frameworks", type=int, default=256, help="c3/c4 catalogue size")
    ap.add_argument("--requests", type=int, default=384, help="c4 requests")
    ap.add_argument("--slots", type=int, default=4, help="c3/c4 requests in flight")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the Widom "
              "insertions shard by index range with no exchange but a 5-block sum",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "c2":
        grid_bench(args)
        return
    if args.workload == "psd":
        print(json.dumps(psd_bench(no_cpu=args.no_cpu)))
        return
    if args.workload == "chain":
        n = min(args.insertions, 1 << 20)
        t0 = time.perf_counter()
        r = chain_bench(no_cpu=args.no_cpu, n=n, steps=args.steps)
        print(json.dumps({
            "metric": METRIC, "value": r["insertions_per_s"], "unit": "insertions/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": round(n / r["insertions_per_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 pair terms, fp64 growth/accumulation", "data": "IRMOF-1 fixture, synthetic trials",
            "config": {"workload": r["workload"], "insertions_per_step": n},
            "cpu_baseline": r.get("cpu_baseline"), "detail": r,
            "wall_s": round(time.perf_counter() - t0, 2)}))
        return
    if args.workload != "c1":
        screening_bench(args)
        return

    import numpy as np
    from raspa2_amd import rhg as R
    from raspa2_amd import workloads as W

    fw = W.c1_framework()
    probe = W.CO2
    n = args.insertions

    # CPU baselines first: forked workers, before this process touches the GPU
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(fw, probe, args.cpu_sample)
    extras = not args.headline_only
    if extras:
        from raspa2_amd import screening as SC
        jobs = SC.catalogue(args.frameworks)
        scr_cpu = screening_cpu(args, jobs)
        grid_ref = grid_cpu_reference(args.grid_check)

    if args.copy_calibration:
        R.stream_copy(512 << 20, reps=3)
    prec = R.PREC_MIXED if args.precision == "mixed" else R.PREC_FP64
    ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=prec)
    _, mol = W.register_probe(ctx, probe)
    info = ctx.info()
    total = (args.warmup + args.steps) * n
    for w in range(args.warmup):
        ctx.widom_run(mol, SEED, total, first=w * n, n=n)
    ctx.synchronize()
    # HIP events time the site-energy launches of the timed steps (the
    # roofline's kernel family) and nothing else: each timed pair of events
    # leaves a gap of some microseconds between launches
    ctx.profile(True, site_energy_only=True)
    ctx.profile_reset()
    ctx.widom_stats(reset=True)
    blocks = R.WidomBlocks()
    ctx.synchronize()
    # each step is enqueued as it would be in a run that reads the block
    # sums at the end: the host prepares step s + 1 while the GPU works on
    # step s, and the sums are collected once, inside the timed region
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.widom_submit(mol, SEED, total, first=(args.warmup + s) * n, n=n)
    ctx.widom_collect(blocks)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = args.steps * n / elapsed
    st = ctx.widom_stats()
    cnt, ms = ctx.profile_get("site_energy")
    se = {"launches": cnt, "total_ms": round(ms, 4), "avg_ms": round(ms / cnt, 5) if cnt else None}
    # the dominant kernel: the orientation-trial list (LJ only for CO2, whose
    # charges enter with the chosen orientation), one launch per step
    cnt_or, ms_or = ctx.profile_get("site_energy:bor")
    se_or = {"launches": cnt_or, "total_ms": round(ms_or, 4),
             "avg_ms": round(ms_or / cnt_or, 5) if cnt_or else None}
    # the other kernel families from one more step, untimed, every family timed
    ctx.profile(True)
    ctx.profile_reset()
    ctx.widom_run(mol, SEED, total + n, first=(args.warmup + args.steps) * n, n=n)
    kernels = {"site_energy": se, "site_energy:bor": se_or}
    for name in ("bin", "fourier", "widom_gen", "widom_select", "block_reduce"):
        cnt, ms = ctx.profile_get(name)
        kernels[name] = {"launches": cnt, "total_ms": round(ms, 4),
                         "avg_ms": round(ms / cnt, 5) if cnt else None,
                         "from": ("one untimed step after the timed ones, every family timed with HIP "
                                  "events: no overlap between families, event gaps included")}
    ctx.profile(False)
    ctx.widom_stats(reset=True)
    # executed pair evaluations of the culled kernels (not the all-pairs count)
    flops = (st["exec_lj_pairs"] * FLOP_LJ_PAIR
             + st["exec_ljcc_pairs"] * (FLOP_LJ_PAIR + FLOP_CC_EXTRA)
             + st["exec_cc_pairs"] * FLOP_CC_PAIR)
    achieved_family = flops / (se["total_ms"] * 1e-3) / 1e12 if se["total_ms"] else None
    # the LJ-only pairs are the orientation-trial list's in this workload
    achieved = (st["exec_lj_pairs"] * FLOP_LJ_PAIR / (se_or["total_ms"] * 1e-3) / 1e12
                if se_or["total_ms"] else None)
    peak = PEAK[args.precision]

    # relative error of this precision against the plain FP64 GPU path on the
    # first timed step, and the plain path's own throughput, timed like the
    # headline (the same warm-up and steps)
    ref_ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_FP64)
    _, mol64 = W.register_probe(ref_ctx, probe)
    b_test = R.WidomBlocks()
    b_ref = R.WidomBlocks()
    ctx.widom_run(mol, SEED, total, first=args.warmup * n, n=n, blocks=b_test)
    for w in range(args.warmup):
        ref_ctx.widom_run(mol64, SEED, total, first=w * n, n=n)
    ref_ctx.synchronize()
    t2 = time.perf_counter()
    ref_ctx.widom_run(mol64, SEED, total, first=args.warmup * n, n=n, blocks=b_ref)
    for s in range(1, args.steps):
        ref_ctx.widom_submit(mol64, SEED, total, first=(args.warmup + s) * n, n=n)
    ref_ctx.widom_collect()
    ref_ctx.synchronize()
    plain_rate = args.steps * n / (time.perf_counter() - t2)
    rho = info["density"]
    e_t = R.widom_estimators(b_test, 300.0, rho)
    e_r = R.widom_estimators(b_ref, 300.0, rho)
    rel_u = abs(e_t["u"] - e_r["u"]) / abs(e_r["u"]) if e_r["u"] else 0.0
    rel_kh = abs(e_t["kh_overall"] - e_r["kh_overall"]) / e_r["kh_overall"]
    est = R.widom_estimators(blocks, 300.0, rho)

    # HBM bytes per site-energy launch, measured by two rocprofv3 --pmc passes
    # (FETCH_SIZE, WRITE_SIZE) over this same command and scaled by the
    # counters' readings for a copy kernel of known size
    traffic, traffic_note = None, "no committed counter summary"
    tpath = os.path.join(ROOT, "profiles", "c1_co2_widom_traffic.json")
    if os.path.exists(tpath) and args.precision == "mixed":
        with open(tpath) as f:
            tj = json.load(f)
        dom = tj["per_kernel"][tj["dominant"]]
        traffic = round(dom["fetch_bytes"] + dom["write_bytes"])
        traffic_note = ("HBM bytes (fetched + written) per launch of " + tj["dominant"] + ", from "
                        + tj["source"])

    cpu_obj = None
    cpu_rel = None
    if cpu is not None:
        cpu_obj = {"value": round(cpu["aggregate"], 2), "unit": "insertions/s",
                   "cores": cpu["cores"], "kind": "port",
                   "single_core": round(cpu["single"], 2),
                   "sample": f"{cpu['sample']} insertions of the same workload (insertions "
                             f"0..{cpu['sample'] - 1} of the job's counter stream), oracle "
                             f"FP64 C restatement, one process per core"}
        # GPU FP64 on the same first insertions as the CPU sample
        m0 = len(cpu["rec0"])
        _, rec_g = ref_ctx.widom_run(mol64, SEED, m0, first=0, n=m0, records=True)
        acc = cpu["rec0"][:, 13] == 4
        flags_equal = bool(np.array_equal(cpu["rec0"][:, 6], rec_g[:, 6]))
        wr = (np.max(np.abs(rec_g[acc, 0] - cpu["rec0"][acc, 0]) / cpu["rec0"][acc, 0])
              if acc.any() else 0.0)
        cpu_rel = {"flags_identical": flags_equal, "max_rel_err_W_fp64": float(wr),
                   "insertions": m0}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "insertions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms, fp64 accumulation/Ewald" if args.precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": "C1 CO2 Widom: rigid 3-site CO2, 4608-atom 51.6 A cubic "
                               "supercell, LJ 12 A shifted + Ewald 1e-6 (6986 k), k_fb=1, "
                               "k_or=10, 300 K",
                   "insertions_per_step": n, "n_atoms": fw.n, "nk": info["nk"],
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": {"bound": "valu",
                     "kernel": "site_energy:bor (k_site_energy_img, orientation trials, shifted LJ)",
                     "achieved": round(achieved, 3),
                     "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": traffic,
                     "traffic_note": traffic_note,
                     "note": "the step's dominant kernel: algorithmic flops of its executed "
                             "(cell-list culled) LJ pair evaluations (29/pair) over its HIP-event "
                             "durations in the timed steps; peak is the published vector rate "
                             "(packed FP32; unpacked issue is half)",
                     "site_energy_family": {
                         "achieved": round(achieved_family, 3), "frac": round(achieved_family / peak, 4),
                         "note": "the three site-energy kernels of a step together: LJ (29/pair), "
                                 "LJ+Coulomb (38) and Coulomb (30) executed pairs over their summed "
                                 "HIP-event durations; the Coulomb kernel of the chosen orientations "
                                 "runs beside the Fourier kernel on a second stream, so its "
                                 "duration includes that sharing of the GPU"}},
        "cpu_baseline": cpu_obj,
        "rel_err_vs_fp64_gpu": {"widom_energy": rel_u, "henry": rel_kh},
        "fp64_plain_gpu": {"value": round(plain_rate, 1), "unit": "insertions/s"},
        "cpu_parity": cpu_rel,
        "estimators": {"henry_mol_kg_pa": est["kh_overall"], "u_gh_K": est["u"],
                       "w": est["w_overall"]},
        "work": {k: float(v) for k, v in st.items()},
        # trial positions (first-bead trials, orientation-trial sites and the
        # Coulomb sites of the selected orientations) and pair evaluations per
        # second of the timed steps; "in range" counts the all-pairs algorithm's
        # pairs, "executed" what the culled kernels evaluated
        "positions_per_s": round((st["fb_sites"] + st["or_sites"] + st["cc_sites"]) / elapsed, 1),
        "pair_evals_per_s": {
            "all_pairs_equivalent": round((st["lj_pairs"] + st["ljcc_pairs"] + st["cc_pairs"]) / elapsed, 1),
            "executed": round((st["exec_lj_pairs"] + st["exec_ljcc_pairs"] + st["exec_cc_pairs"]) / elapsed, 1)},
        "kernels": kernels,
    }
    if extras:
        ctx.close()
        ref_ctx.close()
        # the other half of the metric: batched screening (C3) and serving
        # (C4) with their p50/p99 latency, and the C2 energy grid
        out["screening"] = screening_measure(args, jobs, scr_cpu)
        out["grid"] = grid_measure(grid_ref)
        out["finite_loading"] = loaded_bench(fw, probe)
        out["flexible_chain"] = chain_bench(no_cpu=args.no_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
