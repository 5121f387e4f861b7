"""Widom/CBMC insertion on the GPU against the oracle
(src/mc_moves.c:7844, src/cbmc.c:623-3185)."""
import numpy as np
import pytest

from helpers import gpu_molecule, oracle_molecule, oracle_system, same_bits

pytestmark = pytest.mark.gpu

FIELDS = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
              FBMASK=9, ORMASK=10, W1=11, W2=12, STAGE=13)


def _run_both(W, O, R, fw, probe, charge_method, precision, n, seed, k_fb=None, k_or=None,
              cull=True, images=True):
    sys_o = oracle_system(fw, charge_method=charge_method)
    _, mo = oracle_molecule(sys_o, probe, k_fb, k_or)
    ctx = fw.context(charge_method=charge_method, precision=precision)
    ctx.set_culling(cull, 2.5, images)
    _, mg = gpu_molecule(ctx, probe, k_fb, k_or)
    rec_o = O.widom_run_stream(sys_o, mo, seed, 0, n)
    blocks, rec_g = ctx.widom_run(mg, seed, n, records=True)
    return sys_o, ctx, rec_o, rec_g, blocks


def _check_flags(rec_o, rec_g, check_selection=True):
    for f in ("OVERLAP", "FBMASK", "ORMASK", "STAGE"):
        assert np.array_equal(rec_o[:, FIELDS[f]], rec_g[:, FIELDS[f]]), f
    if check_selection:
        for f in ("SELFB", "SELOR"):
            assert np.array_equal(rec_o[:, FIELDS[f]], rec_g[:, FIELDS[f]]), f


def test_methane_c0_fp64(W, O, R):
    """Single-site LJ probe with 10 first-bead trials: flags, selections and
    U_vdw bit-exact; W to a few ulp (device exp)."""
    fw = W.c0_framework()
    sys_o, ctx, ro, rg, _ = _run_both(W, O, R, fw, W.METHANE, O.CHARGE_NONE, 0, 4000, 11,
                                      k_fb=10)
    _check_flags(ro, rg)
    assert same_bits(ro[:, FIELDS["UVDW"]], rg[:, FIELDS["UVDW"]])
    assert np.allclose(rg[:, 0], ro[:, 0], rtol=1e-13, atol=0)
    assert np.allclose(rg[:, 1], ro[:, 1], rtol=1e-13, atol=0)


@pytest.mark.parametrize("cull", [False, True])
def test_co2_c1_fp64(W, O, R, cull):
    """Rigid CO2 with Ewald in the 4608-atom supercell: first bead + 10
    orientations, LJ biasing, Fourier Delta."""
    fw = W.c1_framework()
    sys_o, ctx, ro, rg, _ = _run_both(W, O, R, fw, W.CO2, O.CHARGE_EWALD, 0, 300, 12, cull=cull)
    _check_flags(ro, rg)
    assert same_bits(ro[:, FIELDS["UVDW"]], rg[:, FIELDS["UVDW"]])
    acc = ro[:, FIELDS["STAGE"]] == 4
    assert acc.sum() > 20
    for f, tol in (("UREAL", 1e-10), ("UFHOST", 1e-9), ("UFADS", 1e-9), ("DU", 1e-10)):
        a, b = ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) <= tol, f
    w_o, w_g = ro[acc, 0], rg[acc, 0]
    assert np.max(np.abs(w_o - w_g) / w_o) <= 1e-9


def test_co2_host_trials_match_device_trials(W, O, R):
    """Host-supplied trials (the caller's RNG) reproduce the device stream."""
    fw = W.c1_framework()
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    _, mo = oracle_molecule(sys_o, W.CO2)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    _, mg = gpu_molecule(ctx, W.CO2)
    n = 128
    fb = np.zeros((n, mo.k_fb, 3))
    cord = np.zeros((n, mo.k_or, mo.n_sites, 3))
    usel = np.zeros((n, 2))
    for i in range(n):
        fb[i], cord[i], usel[i] = O.stream_trials(sys_o, mo, 77, i)
    rec_h = ctx.widom_run_trials(mg, fb, cord, usel)
    _, rec_d = ctx.widom_run(mg, 77, n, records=True)
    assert same_bits(rec_h, rec_d)
    rec_o = np.array([O.widom_insertion_trials(sys_o, mo, fb[i], cord[i], usel[i])
                      for i in range(n)])
    _check_flags(rec_o, rec_h)


@pytest.mark.parametrize("images", [False, True])
def test_mixed_flags_and_tolerance(W, O, R, images):
    """Mixed precision: every overlap flag equals the FP64 reference; W and
    <U> within the stated tolerances."""
    fw = W.c1_framework()
    sys_o, ctx, ro, rg, _ = _run_both(W, O, R, fw, W.CO2, O.CHARGE_EWALD, 1, 300, 13,
                                      images=images)
    for f in ("OVERLAP", "FBMASK", "ORMASK"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    same_sel = (ro[:, FIELDS["SELOR"]] == rg[:, FIELDS["SELOR"]]) & (ro[:, 13] == 4) & (rg[:, 13] == 4)
    assert same_sel.sum() > 20
    w_o, w_g = ro[same_sel, 0], rg[same_sel, 0]
    assert np.max(np.abs(w_o - w_g) / w_o) <= 1e-3
    du_o, du_g = ro[same_sel, 1], rg[same_sel, 1]
    assert np.max(np.abs(du_o - du_g) / np.maximum(np.abs(du_o), 100.0)) <= 1e-5


def test_blocks_and_estimators(W, O, R):
    """Block sums of the device reduction equal the oracle's records summed
    per block; the estimators follow src/statistics.c."""
    fw = W.c0_framework()
    sys_o, ctx, ro, rg, blocks = _run_both(W, O, R, fw, W.METHANE, O.CHARGE_NONE, 0, 5000, 14)
    b = blocks.as_array()
    spc = 20
    bmap = O.block_of_cycles(5000 // spc)
    blk = np.repeat(bmap, spc)
    for k in range(5):
        sel = blk == k
        assert b[k, 0] == sel.sum()
        assert abs(b[k, 1] - ro[sel, 0].sum()) <= 1e-12 * max(1.0, ro[sel, 0].sum())
        swdu = (ro[sel, 0] * ro[sel, 1]).sum()
        assert abs(b[k, 2] - swdu) <= 1e-10 * max(1.0, abs(swdu))
    info = ctx.info()
    est_g = R.widom_estimators(blocks, 300.0, info["density"])
    est_o = O.widom_estimators(b, 300.0, info["density"])
    for key in ("kh", "kh_err", "u", "u_err", "kh_overall", "w_overall"):
        assert abs(est_g[key] - est_o[key]) <= 1e-14 * max(1e-300, abs(est_o[key])), key


def test_chunked_equals_single(W, O, R):
    """A job split into chunks accumulates the same blocks as one call."""
    fw = W.c0_framework()
    ctx = fw.context()
    _, mg = gpu_molecule(ctx, W.METHANE)
    b1, _ = ctx.widom_run(mg, 5, 3000)
    b2 = R.WidomBlocks()
    for first in range(0, 3000, 1000):
        ctx.widom_run(mg, 5, 3000, first=first, n=1000, blocks=b2)
    assert np.allclose(b1.as_array(), b2.as_array(), rtol=1e-14, atol=0)


def _irmof1_helium(O, W):
    import os

    from raspa2_amd import frontend as F
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    share = os.path.join(gold, "share", "raspa")
    ff = F.ForceField.load(os.path.join(share, "forcefield", "GenericMOFs"))
    text = open(os.path.join(share, "structures", "cif", "IRMOF-1.cif")).read()
    fw = F.framework_from_cif(text, ff)
    he = F.read_molecule(os.path.join(share, "molecules", "TraPPE", "helium.def"), ff, k_fb=10)
    assert ff.shifted
    return fw, he, O.CHARGE_NONE


@pytest.mark.parametrize("case", ["helium_irmof1", "co2_c1"])
def test_raspa_mt_order_through_host_trials(W, O, R, case):
    """The drop-in path of a RASPA2 caller that keeps its own random stream:
    the oracle runs RASPA2's MT19937-64 driver (seed 1417341545, the global
    draw order of src/monte_carlo.c:626-649 with src/cbmc.c:652-667,753,1535,
    src/utils.c:688-708,835-877) and records what each insertion drew; the
    same trials through rhg_widom_run_trials give the oracle's records
    insertion by insertion (FP64: flags, masks, selections and stages equal,
    host VDW bit for bit, the other terms and W within 1e-9), and the
    Rosenbluth sum of the records equals the driver's block sums."""
    if case == "helium_irmof1":
        fw, probe, cm = _irmof1_helium(O, W)
        n_cycles, spc = 150, 20
    else:
        fw, probe, cm = W.c1_framework(), W.CO2, O.CHARGE_EWALD
        n_cycles, spc = 30, 20
    print(f"MT-order trials: {case}, {n_cycles * spc} insertions", flush=True)
    sys_o = oracle_system(fw, charge_method=cm)
    _, mo = oracle_molecule(sys_o, probe)
    n, blocks, ro, fb, cord, usel = O.widom_run_mt_trials(sys_o, mo, 1417341545, n_cycles, spc)
    assert n == n_cycles * spc == len(ro)
    ctx = fw.context(charge_method=cm, precision=R.PREC_FP64)
    _, mg = gpu_molecule(ctx, probe)
    rg = ctx.widom_run_trials(mg, fb, cord if probe.n_sites > 1 else None, usel)
    _check_flags(ro, rg)
    assert same_bits(ro[:, FIELDS["UVDW"]], rg[:, FIELDS["UVDW"]])
    acc = ro[:, FIELDS["STAGE"]] == 4
    assert acc.sum() > 50
    for f in ("UREAL", "UFHOST", "UFADS", "DU"):
        a, b = ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) <= 1e-9, f
    assert np.max(np.abs(ro[acc, 0] - rg[acc, 0]) / ro[acc, 0]) <= 1e-9
    assert abs(rg[:, 0].sum() - blocks[:, 1].sum()) <= 1e-9 * blocks[:, 1].sum()
    ctx.close()
