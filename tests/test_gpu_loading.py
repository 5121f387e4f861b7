"""Finite loading on the GPU against the oracle: guest-guest energies
(src/inter_energy.c:321-520), the Fourier Delta of insertion, deletion and
reinsertion against the stored adsorbate sums (src/ewald.c:6859-7565), and
Widom insertion into a loaded framework (src/mc_moves.c:7844-7933)."""
import numpy as np
import pytest

from helpers import gpu_molecule, oracle_molecule, oracle_system, random_positions, same_bits

pytestmark = pytest.mark.gpu

FIELDS = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
              FBMASK=9, ORMASK=10, W1=11, W2=12, STAGE=13, UADSVDW=14, UADSREAL=15)


def _co2(c, v):
    v = np.asarray(v, float)
    v /= np.linalg.norm(v)
    c = np.asarray(c, float)
    return np.array([c + 1.16 * v, c, c - 1.16 * v])


def _load(W, fw, n_mol, seed, types, t_ion=None):
    """n_mol CO2 molecules at random centres (and an ion as the last molecule)."""
    rng = np.random.default_rng(seed)
    start, tys, xyz = [0], [], []
    for p in random_positions(fw.box, n_mol, seed):
        xyz += list(_co2(p, rng.normal(size=3)))
        tys += list(types)
        start.append(start[-1] + 3)
    if t_ion is not None:
        xyz.append(random_positions(fw.box, 1, seed + 1)[0])
        tys.append(t_ion)
        start.append(start[-1] + 1)
    return start, tys, np.array(xyz)


def _both(W, O, R, fw, n_mol, seed, precision=0, ion=True):
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    to, mo = oracle_molecule(sys_o, W.CO2)
    ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=precision)
    tg, mg = gpu_molecule(ctx, W.CO2)
    ion_o = sys_o.add_pseudo_atom(50.0, 3.0, 0.5) if ion else None
    ion_g = ctx.add_pseudo_atom(50.0, 3.0, 0.5) if ion else None
    assert list(to) == list(tg) and ion_o == ion_g
    start, tys, xyz = _load(W, fw, n_mol, seed, tg, ion_g)
    sys_o.set_adsorbates(start, tys, xyz)
    ctx.set_adsorbates(start, tys, xyz)
    return sys_o, ctx, to, tg, mo, mg, ion_g, (start, tys, xyz)


def test_inter_energies(W, O, R):
    """Guest-guest VDW bit for bit (FP64, reference order), real-space
    Coulomb within 1e-12 of its scale, with and without an excluded
    molecule."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, ion, (start, tys, xyz) = _both(W, O, R, fw, 40, 3)
    pts = random_positions(fw.box, 300, 77)
    for t in list(tg) + [ion]:
        for excl in (-1, 7):
            uv, uc = ctx.CalculateInterEnergyAdsorbateAtPosition(pts, t, excl)
            rv = np.array([sys_o.inter_vdw_at(t, p, excl) for p in pts])
            rc = np.array([sys_o.inter_charge_at(t, p, excl) for p in pts])
            assert same_bits(uv, rv), (t, excl)
            assert np.all(np.abs(uc - rc) <= 1e-12 * np.abs(rc).max()), (t, excl)


def test_fourier_moves(W, O, R):
    """Insertion, deletion, reinsertion (CO2 and an ion) against the stored
    adsorbate sums: host and adsorbate changes within 1e-9."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, ion, (start, tys, xyz) = _both(W, O, R, fw, 25, 5)
    new = _co2([9.0, 11.0, 4.0], [0.2, 0.9, -0.3])
    last = len(start) - 2
    cases = [(tg, new, -1), ((), None, 3), (tg, new, 0), ((), None, last),
             ([ion], np.array([[1.0, 2.0, 3.0]]), -1), ([ion], np.array([[1.0, 2.0, 3.0]]), last)]
    for nt, nx, old in cases:
        g = ctx.CalculateEwaldFourierMove(nt, nx, old)
        o = sys_o.fourier_move(nt, nx, old)
        for k in range(2):
            assert abs(g[k] - o[k]) <= 1e-9 * max(1.0, abs(o[k])), (nt, old, k, g[k], o[k])
    # the batched Widom form agrees with the single move at finite loading
    dh, da = ctx.CalculateEwaldFourierAdsorbate(tg, new[None, :, :])
    o = sys_o.fourier_move(tg, new, -1)
    assert abs(dh[0] - o[0]) <= 1e-9 * max(1.0, abs(o[0]))
    assert abs(da[0] - o[1]) <= 1e-9 * max(1.0, abs(o[1]))


@pytest.mark.parametrize("cull", [False, True])
def test_widom_loaded_fp64(W, O, R, cull):
    """CO2 Widom insertion into the loaded supercell (60 CO2 and an ion):
    flags and selections equal the oracle's, host and guest-guest VDW bit for
    bit, the Coulomb, Fourier and W terms within 1e-9."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, ion, _ = _both(W, O, R, fw, 60, 9)
    ctx.set_culling(cull, 2.5)
    ro = O.widom_run_stream(sys_o, mo, 19, 0, 300)
    _, rg = ctx.widom_run(mg, 19, 300, records=True)
    for f in ("OVERLAP", "FBMASK", "ORMASK", "STAGE", "SELFB", "SELOR"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    acc = ro[:, FIELDS["STAGE"]] == 4
    assert acc.sum() > 20
    assert np.any(ro[acc, FIELDS["UADSVDW"]] != 0.0)
    for f in ("UVDW", "UADSVDW"):
        assert same_bits(ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]), f
    for f in ("UREAL", "UADSREAL", "UFHOST", "UFADS", "DU", "W"):
        a, b = ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) <= 1e-9, f


def test_widom_loaded_mixed_and_updates(W, O, R):
    """Mixed precision at finite loading keeps the FP64 flags; replacing the
    adsorbates re-captures the chunk graphs (the records equal a fresh
    context's), and clearing them gives back the infinite-dilution records."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, ion, _ = _both(W, O, R, fw, 60, 13, precision=1)
    ro = O.widom_run_stream(sys_o, mo, 23, 0, 300)
    _, rg = ctx.widom_run(mg, 23, 300, records=True)
    for f in ("OVERLAP", "FBMASK"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    ctx.set_widom_options(chunk=1 << 12, graphs=True)
    _, r1 = ctx.widom_run(mg, 29, 20000, records=True)
    start, tys, xyz = _load(W, fw, 30, 31, tg)
    ctx.set_adsorbates(start, tys, xyz)
    _, r2 = ctx.widom_run(mg, 29, 20000, records=True)
    fresh = fw.context(charge_method=R.CHARGE_EWALD, precision=1)
    _, mf = gpu_molecule(fresh, W.CO2)
    fresh.add_pseudo_atom(50.0, 3.0, 0.5)
    fresh.set_adsorbates(start, tys, xyz)
    fresh.set_widom_options(chunk=1 << 12, graphs=True)
    _, rf = fresh.widom_run(mf, 29, 20000, records=True)
    assert same_bits(r2, rf) and not same_bits(r1, r2)
    ctx.set_adsorbates([0], [], np.zeros((0, 3)))
    clean = fw.context(charge_method=R.CHARGE_EWALD, precision=1)
    _, mc = gpu_molecule(clean, W.CO2)
    clean.set_widom_options(chunk=1 << 12, graphs=True)
    _, rc = clean.widom_run(mc, 29, 20000, records=True)
    _, r3 = ctx.widom_run(mg, 29, 20000, records=True)
    assert same_bits(r3, rc)


def test_retrace_existing_molecules(W, O, R):
    """RetraceMolecule (CBMC_DELETION) of every adsorbed CO2 of a loaded
    supercell, each excluding itself: trial masks, host and guest-guest VDW
    bit for bit, W1, W2 and W within 1e-9 of the oracle."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, ion, (start, tys, xyz) = _both(W, O, R, fw, 40, 21)
    n = len(start) - 2  # the CO2 molecules (the ion is last)
    fbs, cords, olds = [], [], []
    for i in range(n):
        fb, cord, _ = O.stream_trials(sys_o, mo, 47, i)
        fbs.append(fb)
        cords.append(cord)
        olds.append(xyz[start[i]:start[i + 1]])
    rg = np.array([ctx.RetraceMolecule(mg, i, olds[i][None], fbs[i][None], cords[i][None])[0]
                   for i in range(n)])
    rb = ctx.RetraceMolecule(mg, -1, np.array(olds), np.array(fbs), np.array(cords))
    # one batch with a self-exclusion per molecule equals the single calls
    rm = ctx.RetraceMolecule(mg, np.arange(n), np.array(olds), np.array(fbs), np.array(cords))
    assert same_bits(rm, rg)
    ro = np.array([O.widom_retrace_trials(sys_o, mo, i, olds[i], fbs[i], cords[i]) for i in range(n)])
    for f in ("FBMASK", "ORMASK", "STAGE", "SELFB", "SELOR"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    # randomly placed molecules can overlap the framework: their retrace
    # stops (stage 0 or 2) on both sides; the others are compared
    ok = ro[:, FIELDS["STAGE"]] == 4
    assert ok.sum() >= 15
    for f in ("UVDW", "UADSVDW"):
        assert same_bits(ro[ok, FIELDS[f]], rg[ok, FIELDS[f]]), f
    for f in ("W1", "W2", "W", "UREAL", "UADSREAL"):
        a, b = ro[ok, FIELDS[f]], rg[ok, FIELDS[f]]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300)) <= 1e-9, f
    # a configuration that overlaps itself is reported as overlap (stage 0)
    assert np.all(rb[:, FIELDS["STAGE"]] == 0)
    # without excluding themselves the molecules sit on their own atoms
    assert np.all(rb[:, FIELDS["FBMASK"]].astype(np.int64) & 1)


def test_ideal_gas_accumulator_counts_loaded_molecules(W, O, R):
    """WidomIdealGasAccumulated adds Volume / (N + 1) per insertion with the
    N molecules present when the insertions run (src/mc_moves.c:7861-7863):
    100 adsorbed CO2 give V/101 per insertion, cleared adsorbates V/1, and a
    run split into submits gives the same sums as one call."""
    fw = W.c1_framework()
    ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=R.PREC_MIXED)
    tg, mg = gpu_molecule(ctx, W.CO2)
    start, tys, xyz = _load(W, fw, 100, 41, tg)
    ctx.set_adsorbates(start, tys, xyz)
    V = ctx.info()["volume"]
    b, _ = ctx.widom_run(mg, 5, 5000, steps_per_cycle=20)
    a = b.as_array()
    assert a[:, 0].sum() == 5000
    assert np.array_equal(a[:, 3], a[:, 0] * (V / 101.0))
    ctx.widom_submit(mg, 5, 5000, first=0, n=1700)
    ctx.widom_submit(mg, 5, 5000, first=1700, n=3300)
    a2 = ctx.widom_collect().as_array()
    assert np.array_equal(a2[:, 0], a[:, 0])
    assert np.allclose(a2[:, 3], a[:, 3], rtol=1e-15, atol=0)
    ctx.set_adsorbates([0], [], np.zeros((0, 3)))
    a3 = ctx.widom_run(mg, 5, 5000, steps_per_cycle=20)[0].as_array()
    assert np.array_equal(a3[:, 3], a3[:, 0] * V)
    ctx.close()
