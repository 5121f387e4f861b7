"""Asynchronous Widom submission, captured-chunk replay and the host fault
hook (the execution layer around src/mc_moves.c:7844): every form gives the
bitwise block sums of the plain call over the same insertions."""
import numpy as np
import pytest

from helpers import gpu_molecule

pytestmark = pytest.mark.gpu


def _ctx(W, R, fw, charge, chunk, graphs):
    ctx = fw.context(charge_method=charge, precision=R.PREC_MIXED)
    ctx.set_widom_options(chunk=chunk, graphs=graphs)
    return ctx


def test_graph_replay_equals_eager(W, R):
    """Chunks replayed from one captured graph == the same chunks launched
    eagerly (bitwise), CO2 with Ewald; the last chunk is ragged."""
    fw = W.c1_framework(n_rep=1)
    out = []
    for graphs in (False, True):
        ctx = _ctx(W, R, fw, R.CHARGE_EWALD, 1000, graphs)
        _, mg = gpu_molecule(ctx, W.CO2)
        b, _ = ctx.widom_run(mg, 31, 5500)
        st = ctx.widom_stats()
        out.append((b.as_array(), st))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[1][1]["insertions"] == 5500
    for k in ("fb_sites", "or_sites", "cc_sites", "fourier_mols", "exec_lj_pairs",
              "exec_ljcc_pairs", "exec_cc_pairs"):
        assert out[0][1][k] == out[1][1][k], k


def test_graph_replay_repeated_calls(W, R):
    """A cached graph is replayed by later calls with other ranges: each call
    equals a fresh eager context's result."""
    fw = W.c0_framework()
    g = _ctx(W, R, fw, R.CHARGE_NONE, 512, True)
    e = _ctx(W, R, fw, R.CHARGE_NONE, 512, False)
    _, mg = gpu_molecule(g, W.METHANE)
    _, me = gpu_molecule(e, W.METHANE)
    for first, n in ((0, 3000), (3000, 2048), (5048, 4952)):
        bg, _ = g.widom_run(mg, 9, 10000, first=first, n=n)
        be, _ = e.widom_run(me, 9, 10000, first=first, n=n)
        assert np.array_equal(bg.as_array(), be.as_array()), (first, n)


def test_submit_collect_equals_run(W, R):
    """Two contexts in flight at once; each collect equals its plain run."""
    fws = [W.c0_framework(), W.c1_framework(n_rep=1)]
    ref, ctxs, mols = [], [], []
    for fw, probe, ch in ((fws[0], W.METHANE, R.CHARGE_NONE), (fws[1], W.CO2, R.CHARGE_EWALD)):
        c = _ctx(W, R, fw, ch, 4096, True)
        _, m = gpu_molecule(c, probe)
        ref.append(c.widom_run(m, 17, 20000)[0].as_array())
        ctxs.append(c)
        mols.append(m)
    for c, m in zip(ctxs, mols):
        c.widom_submit(m, 17, 20000, first=0, n=8192)
        c.widom_submit(m, 17, 20000, first=8192, n=11808)
    got = [c.widom_collect().as_array() for c in ctxs]
    for k in range(2):
        assert np.allclose(got[k], ref[k], rtol=1e-14, atol=0)
        assert ctxs[k].widom_query()


def test_fault_hook_and_resume(W, R):
    """The host hook fails a chunk before launching it; resuming from the
    last whole chunk gives the uninterrupted result."""
    fw = W.c0_framework()
    ctx = _ctx(W, R, fw, R.CHARGE_NONE, 1000, True)
    _, m = gpu_molecule(ctx, W.METHANE)
    ref, _ = ctx.widom_run(m, 23, 6000, first=0, n=6000)
    ctx.widom_collect()
    blocks = R.WidomBlocks()
    ctx.inject_fault(3)
    with pytest.raises(R.RhgError) as ei:
        ctx.widom_submit(m, 23, 6000, first=0, n=6000)
    assert ei.value.code == R.ERR_FAULT_INJECTED
    # the three chunks before the fault completed; their sums are collected
    ctx.widom_collect(blocks)
    assert np.sum(blocks.as_array()[:, 0]) == 3000
    ctx.widom_submit(m, 23, 6000, first=3000, n=3000)
    ctx.widom_collect(blocks)
    assert np.allclose(blocks.as_array(), ref.as_array(), rtol=1e-14, atol=0)
