"""Screening service (rhg_server_*): continuous batching over resident
frameworks.  Each request's estimators are bitwise the same alone as in the
mix, after a fault-and-retry, and under a small HBM budget that forces
evictions; results agree with the plain Widom call (src/mc_moves.c:7844)."""
import numpy as np
import pytest

from helpers import gpu_molecule

pytestmark = pytest.mark.gpu

SLICE = 8192


@pytest.fixture(scope="module")
def jobs(W):
    out = []
    for i in range(6):
        j = W.c3_job(i)
        j.n_trials = min(j.n_trials, 20000)  # test-sized requests
        out.append(j)
    return out


def _serve(S, R, jobs, idx, slots=3, arrivals=None, budget=0, faults=()):
    srv = S.Server(slots=slots, precision=R.PREC_MIXED, slice_=SLICE, hbm_budget=budget)
    for k in idx:
        srv.add_framework(k, jobs[k].framework)
    rids = {}
    for n, k in enumerate(idx):
        a = 0.0 if arrivals is None else arrivals[n]
        rids[k] = srv.submit(k, jobs[k].probe, jobs[k].seed, jobs[k].n_trials, arrival_s=a)
    for k, after in faults:
        srv.inject_fault(rids[k], after)
    srv.run()
    res = {k: srv.result(rids[k]) for k in idx}
    st = srv.stats()
    srv.close()
    return res, st


@pytest.fixture(scope="module")
def S():
    from raspa2_amd import serve
    return serve


def test_alone_equals_mix(W, R, S, jobs):
    mix, st = _serve(S, R, jobs, range(6), slots=3,
                     arrivals=np.cumsum(np.random.default_rng(5).exponential(0.002, 6)))
    assert st["created"] == 6
    for k in range(6):
        alone, _ = _serve(S, R, jobs, [k], slots=1)
        a, m = alone[k], mix[k]
        assert a["status"] == 0 and m["status"] == 0
        assert np.array_equal(a["blocks"].as_array(), m["blocks"].as_array()), k
        assert m["latency_s"] > 0


def test_matches_plain_call(W, R, S, jobs):
    """The sliced service result equals one plain rhg_widom_run (the sums are
    grouped differently, so to rounding)."""
    k = 1
    res, _ = _serve(S, R, jobs, [k], slots=1)
    j = jobs[k]
    ctx = j.framework.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
    _, m = gpu_molecule(ctx, j.probe)
    b, _ = ctx.widom_run(m, j.seed, j.n_trials)
    assert np.allclose(res[k]["blocks"].as_array(), b.as_array(), rtol=1e-12, atol=0)


def test_fault_retry_bitwise(W, R, S, jobs):
    """A slice failed by the host hook is discarded and run again: the result
    is bitwise the uninterrupted one."""
    ref, _ = _serve(S, R, jobs, [0, 2], slots=2)
    got, st = _serve(S, R, jobs, [0, 2], slots=2, faults=[(2, 1)])
    assert st["retries"] == 1
    assert got[2]["retries"] == 1
    for k in (0, 2):
        assert got[k]["status"] == 0
        assert np.array_equal(got[k]["blocks"].as_array(), ref[k]["blocks"].as_array())


def test_small_budget_evicts(W, R, S, jobs):
    """With room for about two frameworks, contexts are evicted and rebuilt;
    every result is unchanged."""
    est = [S.estimate_footprint(jobs[k].framework, jobs[k].probe, chunk=SLICE) for k in range(6)]
    budget = int(2.2 * max(est))
    ref, _ = _serve(S, R, jobs, range(6), slots=2)
    got, st = _serve(S, R, jobs, list(range(6)), slots=2, budget=budget)
    assert st["peak_resident"] <= 1.3 * budget  # admission by estimate, then actual bytes
    assert st["evicted"] >= 1
    for k in range(6):
        assert got[k]["status"] == 0
        assert np.array_equal(got[k]["blocks"].as_array(), ref[k]["blocks"].as_array())


def test_too_large_is_rejected(W, R, S, jobs):
    est = S.estimate_footprint(jobs[0].framework, jobs[0].probe, chunk=SLICE)
    got, _ = _serve(S, R, jobs, [0], slots=1, budget=est // 2)
    assert got[0]["status"] == R.ERR_NOT_ADMITTED


def test_estimate_against_actual(W, R, S, jobs):
    """The admission estimate is within [0.6, 1.6] of what a context holds
    after a run with the same chunk."""
    for k in (0, 1, 3):
        j = jobs[k]
        ctx = j.framework.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
        ctx.set_widom_options(chunk=SLICE, graphs=True)
        _, m = gpu_molecule(ctx, j.probe)
        ctx.widom_run(m, j.seed, 3 * SLICE)
        actual = ctx.footprint()
        est = S.estimate_footprint(j.framework, j.probe, chunk=SLICE)
        assert 0.6 <= est / actual <= 1.6, (k, est, actual)


def test_scratch_reused_by_growing_cell_grids(W, R, S):
    """One slot serves frameworks whose cell grids grow from request to
    request, so each reuses the slot's binning scratch sized (with headroom)
    by an earlier, smaller one; every request's block sums equal the same
    request served alone on a fresh service (a regression test: the per-cell
    counts must be cleared beyond the cells of the request that allocated
    them)."""
    cand = []
    for i in range(40):
        j = W.c3_job(i)
        j.n_trials = 4096
        cand.append(j)
    # frameworks by increasing cell volume (the cell grid grows with it)
    vol = [float(np.prod(np.diag(np.asarray(j.framework.box).reshape(3, 3)))) for j in cand]
    order = list(np.argsort(vol))[::3][:10]
    jobs = {int(k): cand[int(k)] for k in order}
    seq, _ = _serve(S, R, jobs, list(jobs), slots=1)
    for k in jobs:
        alone, _ = _serve(S, R, jobs, [k], slots=1)
        assert seq[k]["status"] == 0 and alone[k]["status"] == 0
        assert np.array_equal(seq[k]["blocks"].as_array(), alone[k]["blocks"].as_array()), k
