"""Screening service over librhg (rhg_server_*): a resident framework cache
with HBM admission and continuous batching of Widom requests on one GPU.

The reference answers one request per process (python/raspa2.py:97-110);
this is the many-requests form of the same Widom move
(src/mc_moves.c:7844), with each request's result independent of what runs
beside it.
"""
import ctypes

import numpy as np

from . import rhg as R

_P, _D, _I32, _I64, _U64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int64, \
    ctypes.c_uint64
MS = R.MAX_SITES


class ProbeT(ctypes.Structure):
    _fields_ = [("n_sites", _I32), ("eps_k", _D * MS), ("sigma", _D * MS), ("charge", _D * MS),
                ("body", (_D * 3) * MS), ("start", _I32), ("k_fb", _I32), ("k_or", _I32)]


class RequestT(ctypes.Structure):
    _fields_ = [("framework", _I64), ("probe", ProbeT), ("seed", _U64), ("n_insertions", _I64),
                ("steps_per_cycle", _I32), ("arrival_s", _D)]


class ResultT(ctypes.Structure):
    _fields_ = [("status", _I32), ("retries", _I32), ("blocks", R.WidomBlocks),
                ("queued_s", _D), ("started_s", _D), ("finished_s", _D), ("latency_s", _D)]


def _decl():
    L = R.lib

    def f(name, *args, res=ctypes.c_int):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("rhg_server_create", ctypes.c_int, ctypes.c_size_t, _I32, _I32, _I64, ctypes.POINTER(_P))
    f("rhg_server_destroy", _P, res=None)
    f("rhg_server_last_error", _P, res=ctypes.c_char_p)
    f("rhg_server_add_framework", _P, _I64, ctypes.POINTER(R.Framework), ctypes.POINTER(R.CellT),
      ctypes.POINTER(R.ForceField))
    f("rhg_server_set_tuning", _P, _I32, _P, _P, _P, _P, _P)
    f("rhg_server_use_grids", _P, _I64, _D)
    f("rhg_server_submit", _P, ctypes.POINTER(RequestT), ctypes.POINTER(_I64))
    f("rhg_server_inject_fault", _P, _I64, _I64)
    f("rhg_server_run", _P)
    f("rhg_server_start", _P)
    f("rhg_server_stop", _P)
    f("rhg_server_wait", _P, _I64, _D)
    f("rhg_server_result", _P, _I64, ctypes.POINTER(ResultT))
    f("rhg_server_stats", _P, _P)
    f("rhg_estimate_footprint", ctypes.POINTER(R.Framework), ctypes.POINTER(R.CellT),
      ctypes.POINTER(R.ForceField), _I32, _I64, _I32, _I32, _I32, _I32, _D, _I32,
      res=ctypes.c_size_t)


_decl()


def _structs(spec, charge_method, cutoff=12.0, ewald_precision=1e-6, temperature=300.0):
    """ctypes framework/cell/force-field of a workloads.FrameworkSpec; the
    arrays are returned too so that they outlive the call."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (spec.xyz[:, 0], spec.xyz[:, 1], spec.xyz[:, 2], spec.charge, spec.eps_k, spec.sigma,
             spec.mass)]
    p = [a.ctypes.data_as(ctypes.POINTER(_D)) for a in arrs]
    fw = R.Framework(len(arrs[0]), p[0], p[1], p[2], p[3], p[4], p[5], p[6])
    cell = R.CellT()
    for i, v in enumerate(np.asarray(spec.box, dtype=np.float64).reshape(9)):
        cell.box[i] = v
    cell.boundary = spec.bc
    ff = R.ForceField(cutoff, 1, charge_method, cutoff, ewald_precision, temperature)
    return fw, cell, ff, arrs


def probe_struct(probe):
    p = ProbeT()
    p.n_sites = probe.n_sites
    for j, (e, s, q) in enumerate(probe.sites):
        p.eps_k[j], p.sigma[j], p.charge[j] = e, s, q
        for d in range(3):
            p.body[j][d] = float(probe.body[j][d])
    p.start, p.k_fb, p.k_or = probe.start, probe.k_fb, probe.k_or
    return p


def estimate_footprint(spec, probe, charge_method=R.CHARGE_EWALD, chunk=1 << 18, cell_size=2.5,
                       precision=R.PREC_MIXED):
    """Device bytes of a context of `spec` for `probe` (host arithmetic only)."""
    fw, cell, ff, _keep = _structs(spec, charge_method)
    n_types = len(set(probe.sites))
    n_charged = sum(1 for _, _, q in probe.sites if abs(q) > 1e-10)
    return R.lib.rhg_estimate_footprint(ctypes.byref(fw), ctypes.byref(cell), ctypes.byref(ff),
                                        n_types, chunk, probe.n_sites, n_charged, probe.k_fb,
                                        probe.k_or, cell_size, precision)


class Server:
    """rhg_server: see include/rhg.h (screening service)."""

    def __init__(self, device=0, hbm_budget=0, slots=4, precision=R.PREC_MIXED, slice_=1 << 18):
        h = _P()
        rc = R.lib.rhg_server_create(device, hbm_budget, slots, precision, slice_, ctypes.byref(h))
        if rc:
            raise R.RhgError(f"rhg_server_create: status {rc}", rc)
        self.h = h
        self.temperature = {}
        self._n = 0

    def close(self):
        if getattr(self, "h", None):
            R.lib.rhg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = R.lib.rhg_server_last_error(self.h)
            raise R.RhgError(f"{what}: status {rc}: {msg.decode() if msg else ''}", rc)

    def add_framework(self, fid, spec, charge_method=R.CHARGE_EWALD, temperature=300.0):
        fw, cell, ff, _keep = _structs(spec, charge_method, temperature=temperature)
        self._check(R.lib.rhg_server_add_framework(self.h, fid, ctypes.byref(fw), ctypes.byref(cell),
                                                   ctypes.byref(ff)), "add_framework")
        self.temperature[fid] = temperature

    def set_tuning(self, rows):
        """rows: [(max_atoms, cell_size, chunk[, sites_per_lane, waves_per_block])]"""
        ma = np.array([r[0] for r in rows], dtype=np.int32)
        cs = np.array([r[1] for r in rows], dtype=np.float64)
        ch = np.array([r[2] for r in rows], dtype=np.int64)
        la = np.array([r[3] if len(r) > 3 else 4 for r in rows], dtype=np.int32)
        wa = np.array([r[4] if len(r) > 4 else 4 for r in rows], dtype=np.int32)
        self._check(R.lib.rhg_server_set_tuning(self.h, len(rows), ma.ctypes.data_as(_P),
                                                cs.ctypes.data_as(_P), ch.ctypes.data_as(_P),
                                                la.ctypes.data_as(_P), wa.ctypes.data_as(_P)),
                    "set_tuning")

    def use_grids(self, fid, spacing):
        """Serve framework fid's requests from energy grids at `spacing` [A] (0: direct sums)."""
        self._check(R.lib.rhg_server_use_grids(self.h, fid, spacing), "use_grids")

    def submit(self, fid, probe, seed, n_insertions, steps_per_cycle=20, arrival_s=0.0):
        r = RequestT(fid, probe_struct(probe), seed, n_insertions, steps_per_cycle, arrival_s)
        i = _I64()
        self._check(R.lib.rhg_server_submit(self.h, ctypes.byref(r), ctypes.byref(i)), "submit")
        return i.value

    def inject_fault(self, rid, after_chunks):
        self._check(R.lib.rhg_server_inject_fault(self.h, rid, after_chunks), "inject_fault")

    def run(self):
        """Serves every submitted request (and those submitted meanwhile), then returns."""
        self._check(R.lib.rhg_server_run(self.h), "run")

    def start(self):
        """Starts the request loop; submit() may be called from any thread meanwhile."""
        self._check(R.lib.rhg_server_start(self.h), "start")

    def stop(self):
        """Finishes every submitted request and stops the loop."""
        self._check(R.lib.rhg_server_stop(self.h), "stop")

    def wait(self, rid, timeout_s=-1.0):
        """True once request rid has finished (False when timeout_s passed first)."""
        rc = R.lib.rhg_server_wait(self.h, rid, timeout_s)
        if rc < 0:
            self._check(rc, "wait")
        return rc == 1

    def result(self, rid):
        r = ResultT()
        self._check(R.lib.rhg_server_result(self.h, rid, ctypes.byref(r)), "result")
        return dict(status=r.status, retries=r.retries, blocks=r.blocks, queued_s=r.queued_s,
                    started_s=r.started_s, finished_s=r.finished_s, latency_s=r.latency_s)

    def stats(self):
        out = np.zeros(20)
        self._check(R.lib.rhg_server_stats(self.h, out.ctypes.data_as(_P)), "stats")
        keys = ("created", "evicted", "peak_resident", "resident", "retries", "budget",
                "idle_wait_s", "requests", "setup_s", "submit_s", "collect_s", "grids_built",
                "grids_evicted", "grid_bytes", "finished", "insertions", "exec_lj_pairs",
                "exec_ljcc_pairs", "exec_cc_pairs")
        return dict(zip(keys, out))
