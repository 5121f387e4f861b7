"""Flexible-chain CBMC Widom insertion on the GPU against the oracle
(WidomAdsorbateMove src/mc_moves.c:7844-7946 with GrowMolecule, HandleFirstBead,
GenerateTrialOrientationsSimpleSphere/MCScheme, ComputeExternalEnergies and
Rosen, src/cbmc.c:623-795, 1154-2482, 2762-2849, 3056-3185): both sides draw
from the same per-insertion counter sub-streams, so every insertion is
compared on its own -- stage, Rosenbluth weight, energies and the grown
molecule."""
import math

import numpy as np
import pytest

from helpers import gpu_chain, irmof1_dubbeldam, oracle_chain, oracle_system

pytestmark = pytest.mark.gpu


def _empty_pair(O, R, L=30.0):
    box = np.diag([L, L, L]).T.reshape(-1)
    s = O.System(box, O.BC_RECTANGULAR, [], [], [], [], [], [], cutoff_vdw=12.0, temperature=300.0)
    e = np.zeros(0)
    ctx = R.Context(box, R.BC_RECTANGULAR, e, e, e, e, e, charge=e, mass=e, cutoff_vdw=12.0,
                    shifted=True, temperature=300.0, precision=R.PREC_FP64)
    return s, ctx


def _compare(O, R, rec_o, xyz_o, rec_g, xyz_g, rtol=1e-9):
    stage_o, stage_g = rec_o[:, O.CREC["STAGE"]], rec_g[:, R.CREC["STAGE"]]
    bad = np.flatnonzero(stage_o != stage_g)
    assert len(bad) == 0, f"stages differ at {bad[:10]}: {stage_o[bad[:10]]} vs {stage_g[bad[:10]]}"
    assert np.array_equal(rec_o[:, O.CREC["STEP"]], rec_g[:, R.CREC["STEP"]])
    for f in ("W", "DU", "UHOSTVDW", "UINTRA", "UBOND", "UBEND", "UTORSION", "W1"):
        a, b = rec_o[:, O.CREC[f]], rec_g[:, R.CREC[f]]
        scale = np.maximum(np.abs(a), 1e-300 if f in ("W", "W1") else 1.0)
        err = np.abs(a - b) / scale
        assert err.max() <= rtol, f"{f}: worst {err.max():.3g} at {int(np.argmax(err))}"
    ok = stage_o >= 4
    assert np.allclose(xyz_o[ok], xyz_g[ok], rtol=0, atol=1e-9)


def test_chain_empty_box(O, W, R):
    """Butane and hexane without a framework: internal walk, torsion
    selection and intra-molecular LJ only; W against the torsion quadrature."""
    s, ctx = _empty_pair(O, R)
    for n, k in ((4, 10), (6, 10)):
        probe = W.trappe_alkane(n)
        _, co = oracle_chain(s, probe)
        _, cg = gpu_chain(ctx, probe)
        rec_o, xyz_o = O.chain_run_stream(s, co, 17, 0, 300, positions=True)
        _, rec_g, xyz_g = ctx.widom_chain_run(cg, 17, 300, records=True, positions=True)
        print(f"C{n}: accepted {int((rec_g[:, R.CREC['STAGE']] == 4).sum())}/300", flush=True)
        _compare(O, R, rec_o, xyz_o, rec_g, xyz_g)
    # many insertions: <W> of butane is the mean of exp(-beta U_torsion)
    probe = W.trappe_alkane(4, k_trial=10, k_torsion=10)
    _, cg = gpu_chain(ctx, probe)
    _, rec, _ = ctx.widom_chain_run(cg, 5, 200000, records=True)
    w = rec[:, R.CREC["W"]]
    phi = np.linspace(-math.pi, math.pi, 400001)
    u = 0.0 + 355.03 * (1 + np.cos(phi)) - 68.19 * (1 - np.cos(2 * phi)) + 791.32 * (1 + np.cos(3 * phi))
    wq = np.exp(-u / 300.0).mean()
    assert abs(w.mean() - wq) < 4.0 * w.std() / math.sqrt(len(w))


def test_chain_irmof1_fp64(O, W, R):
    """Pentane and hexane in IRMOF-1 (FP64): every insertion as the oracle's."""
    fw = irmof1_dubbeldam()[0]
    s = oracle_system(fw)
    ctx = fw.context(precision=R.PREC_FP64)
    for n in (5, 6):
        probe = W.trappe_alkane(n)
        _, co = oracle_chain(s, probe)
        _, cg = gpu_chain(ctx, probe)
        rec_o, xyz_o = O.chain_run_stream(s, co, 23, 100, 200, positions=True)
        blocks, rec_g, xyz_g = ctx.widom_chain_run(cg, 23, 300, first=100, n=200, records=True,
                                                   positions=True)
        print(f"C{n}: stages {np.unique(rec_g[:, R.CREC['STAGE']], return_counts=True)}", flush=True)
        _compare(O, R, rec_o, xyz_o, rec_g, xyz_g)
        # the block sums hold the records' W and W dU
        b = blocks.as_array()
        assert b[:, 0].sum() == 200
        assert b[:, 1].sum() == pytest.approx(rec_g[:, R.CREC["W"]].sum(), rel=1e-12)
        assert b[:, 2].sum() == pytest.approx((rec_g[:, R.CREC["W"]] * rec_g[:, R.CREC["DU"]]).sum(),
                                              rel=1e-10)


def test_chain_started_in_the_middle_and_pockets(O, W, R):
    """Hexane grown from bead 2 (steps that grow away from the first bead in
    both directions), with blocked pockets rejecting part of the grown
    molecules."""
    fw = irmof1_dubbeldam()[0]
    s = oracle_system(fw)
    ctx = fw.context(precision=R.PREC_FP64)
    ub = np.asarray(fw.unit_box)
    centres = np.array([[6.0, 6.0, 6.0], [19.0, 19.0, 19.0]])
    radii = np.array([5.0, 5.0])
    s.set_block_pockets(centres, radii, ub)
    ctx.set_block_pockets(centres, radii, ub)
    probe = W.trappe_alkane(6, start=2)
    _, co = oracle_chain(s, probe)
    _, cg = gpu_chain(ctx, probe)
    rec_o, xyz_o = O.chain_run_stream(s, co, 29, 0, 200, positions=True)
    _, rec_g, xyz_g = ctx.widom_chain_run(cg, 29, 200, records=True, positions=True)
    stages = rec_g[:, R.CREC["STAGE"]]
    print(f"stages {np.unique(stages, return_counts=True)}", flush=True)
    assert (stages == 5).any() and (stages == 4).any()
    _compare(O, R, rec_o, xyz_o, rec_g, xyz_g)


def test_chain_mixed_precision(O, W, R):
    """Mixed precision (FP32 pair terms): the same insertions within the
    pair-term rounding wherever both sides select the same trials."""
    fw = irmof1_dubbeldam()[0]
    s = oracle_system(fw)
    ctx = fw.context(precision=R.PREC_MIXED)
    probe = W.trappe_alkane(5)
    _, co = oracle_chain(s, probe)
    _, cg = gpu_chain(ctx, probe)
    rec_o = O.chain_run_stream(s, co, 31, 0, 400)
    _, rec_g, _ = ctx.widom_chain_run(cg, 31, 400, records=True)
    same = rec_o[:, O.CREC["STAGE"]] == rec_g[:, R.CREC["STAGE"]]
    assert same.mean() > 0.97
    ok = same & (rec_o[:, O.CREC["STAGE"]] == 4)
    close = np.isclose(rec_o[ok, O.CREC["UHOSTVDW"]], rec_g[ok, R.CREC["UHOSTVDW"]], rtol=1e-4, atol=1e-2)
    assert close.mean() > 0.95
    wo, wg = rec_o[:, O.CREC["W"]].sum(), rec_g[:, R.CREC["W"]].sum()
    assert wg == pytest.approx(wo, rel=0.05)


def test_chain_errors(W, R):
    fw = irmof1_dubbeldam()[0]
    ctx = fw.context(precision=R.PREC_FP64)
    probe = W.trappe_alkane(4)
    _, cg = gpu_chain(ctx, probe, k_trial=20, k_torsion=10)
    with pytest.raises(R.RhgError) as e:
        ctx.widom_chain_run(cg, 1, 10)
    assert e.value.code == R.ERR_INVALID
    charged = W.trappe_alkane(4)
    charged.sites = [(98.0, 3.75, 0.1)] + charged.sites[1:]
    _, cg = gpu_chain(ctx, charged)
    with pytest.raises(R.RhgError) as e:
        ctx.widom_chain_run(cg, 1, 10)
    assert e.value.code == R.ERR_UNSUPPORTED
    broken = W.trappe_alkane(4)
    broken.bonds = broken.bonds[:2]  # atom 3 not connected
    _, cg = gpu_chain(ctx, broken)
    with pytest.raises(R.RhgError) as e:
        ctx.widom_chain_run(cg, 1, 10)
    assert e.value.code == R.ERR_INVALID
    # the context still works after the rejected calls
    _, cg = gpu_chain(ctx, probe)
    blocks, rec, _ = ctx.widom_chain_run(cg, 1, 50, records=True)
    assert blocks.as_array()[:, 0].sum() == 50


def test_ideal_rosenbluth_weights_printed(O, R):
    """The reference's IdealRosenbluthWeights-C5-C9-573K example on the
    device: 10^6 growths of each n-alkane in an empty box at 573 K give the
    printed average Rosenbluth weights within three combined standard errors
    (fixture tests/golden/ideal_rosenbluth_c5_c9.json; the run's printed
    torsion parameters)."""
    import json
    import os
    from raspa2_amd import frontend as F
    here = os.path.dirname(os.path.abspath(__file__))
    G = json.load(open(os.path.join(here, "golden", "ideal_rosenbluth_c5_c9.json")))
    share = os.path.join(here, "golden", "share", "raspa")
    ff = F.ForceField.load(os.path.join(share, "forcefield", "GarciaPerez2006"))
    box = np.diag(G["box"]).T.reshape(-1)
    e = np.zeros(0)
    ctx = R.Context(box, R.BC_RECTANGULAR, e, e, e, e, e, charge=e, mass=e, cutoff_vdw=12.0,
                    shifted=ff.shifted, temperature=G["temperature"], precision=R.PREC_FP64)
    n = 1_000_000
    for name in ("C5", "C6", "C7", "C8", "C9"):
        p = F.read_molecule(os.path.join(share, "molecules", "TraPPE", name + ".def"), ff)
        tp = G["printed_torsions"][name]["p"]
        types = [ctx.add_pseudo_atom(eps, sg, q) for eps, sg, q in p.sites]
        ch = R.Chain(types, p.bonds, p.bends, [t[:5] + tuple(tp) for t in p.torsions], p.intra)
        _, rec, _ = ctx.widom_chain_run(ch, 7, n, records=True)
        w = rec[:, R.CREC["W"]]
        ref = G["average_widom"][name]
        sig = math.hypot(w.std() / math.sqrt(n), ref["plus_minus"] / 2.0)
        print(f"{name}: {w.mean():.6g} vs printed {ref['average']:.6g} +/- {ref['plus_minus']:.2g}", flush=True)
        assert abs(w.mean() - ref["average"]) < 3.0 * sig, (name, w.mean(), ref)
