"""UseTabularGrid on the GPU: energies at positions and Widom insertion
served from device-built grids, against the oracle running on the same grid
data (src/framework_energy.c:73-76, 197-200; src/grids.c:979-1077,
1200-1295)."""
import numpy as np
import pytest

from helpers import gpu_molecule, oracle_molecule, oracle_system, random_positions, same_bits

pytestmark = pytest.mark.gpu

FIELDS = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
              FBMASK=9, ORMASK=10, W1=11, W2=12, STAGE=13)


def _grids(W, O, R, fw, probe, spacing, vdw=True, coulomb=True):
    """Context and oracle with the probe's pseudo-atoms and, in use on both,
    one VDW grid per distinct site and the Coulomb grid."""
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    to, mo = oracle_molecule(sys_o, probe)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    tg, mg = gpu_molecule(ctx, probe)
    geom = O.grid_geometry(fw.unit_box, spacing)
    keep = []
    if vdw:
        made = {}
        for (e, s, q), a, b in zip(probe.sites, to, tg):
            if (e, s) not in made:
                g = ctx.MakeGrid(b, fw.unit_box, spacing)
                assert np.array_equal(g.geom, geom)
                made[(e, s)] = (g, g.download())
            g, data = made[(e, s)]
            ctx.use_grid(b, g)
            sys_o.use_grid(a, data, geom, fw.unit_box)
        keep += [v[0] for v in made.values()]
    if coulomb:
        gc = ctx.MakeCoulombGrid(fw.unit_box, spacing)
        ctx.use_coulomb_grid(gc)
        sys_o.use_coulomb_grid(gc.download(), geom, fw.unit_box)
        keep.append(gc)
    return sys_o, ctx, to, tg, mo, mg, keep


def test_energies_from_grids(W, O, R):
    """Energies at positions of every CO2 site from the grids equal the
    oracle's interpolation of the same grids (1e-9 relative), overlap flags
    included; the Fourier path is unaffected."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, keep = _grids(W, O, R, fw, W.CO2, 0.5)
    pts = random_positions(fw.box, 1500, 61)
    for a, b in zip(to, tg):
        uv, uc, ov = ctx.energy_at_positions(pts, b)
        ref_v = np.array([sys_o.vdw_at(a, p) for p in pts])
        ref_c = np.array([sys_o.charge_at(a, p) for p in pts])
        assert np.all(np.abs(uv - ref_v) <= 1e-9 * np.maximum(1.0, np.abs(ref_v)))
        assert np.all(np.abs(uc - ref_c) <= 1e-9 * np.maximum(1.0, np.abs(ref_c)))
        assert np.array_equal(ov, (ref_v >= 1e7) | (ref_c >= 1e7))
        assert ov.any() and not ov.all()


def test_one_term_from_grid(W, O, R):
    """A VDW grid alone leaves the Coulomb term on the pair kernels (bit for
    bit the gridless values) and vice versa; detaching restores both."""
    fw = W.c1_framework()
    ctx0 = fw.context(charge_method=R.CHARGE_EWALD)
    t0 = ctx0.add_pseudo_atom(79.0, 3.05, -0.35)
    pts = random_positions(fw.box, 800, 67)
    uv0, uc0, ov0 = ctx0.energy_at_positions(pts, t0)
    sys_o, ctx, to, tg, mo, mg, keep = _grids(W, O, R, fw, W.CO2, 0.5, coulomb=False)
    uv, uc, _ = ctx.energy_at_positions(pts, tg[0])
    assert same_bits(uc, uc0) and not same_bits(uv, uv0)
    ctx.use_grid(tg[0], None)
    gc = ctx.MakeCoulombGrid(fw.unit_box, 0.5)
    ctx.use_coulomb_grid(gc)
    uv, uc, _ = ctx.energy_at_positions(pts, tg[0])
    assert same_bits(uv, uv0) and not same_bits(uc, uc0)
    ctx.use_coulomb_grid(None)
    uv, uc, ov = ctx.energy_at_positions(pts, tg[0])
    assert same_bits(uv, uv0) and same_bits(uc, uc0) and np.array_equal(ov, ov0)
    # freeing a grid in use detaches it first
    ctx.use_coulomb_grid(gc)
    gc.close()
    uv, uc, _ = ctx.energy_at_positions(pts, tg[0])
    assert same_bits(uc, uc0)


@pytest.mark.parametrize("cull", [False, True])
def test_widom_from_grids(W, O, R, cull):
    """CO2 Widom insertion with VDW and Coulomb grids in use: flags and
    selections equal the oracle's on the same grids, energies and W within
    1e-9 (the Fourier term is still the Ewald sum)."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg, mo, mg, keep = _grids(W, O, R, fw, W.CO2, 0.5)
    ctx.set_culling(cull, 2.5)
    ro = O.widom_run_stream(sys_o, mo, 17, 0, 300)
    _, rg = ctx.widom_run(mg, 17, 300, records=True)
    for f in ("OVERLAP", "FBMASK", "ORMASK", "STAGE", "SELFB", "SELOR"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    acc = ro[:, FIELDS["STAGE"]] == 4
    assert acc.sum() > 20
    for f in ("UVDW", "UREAL", "UFHOST", "UFADS", "DU", "W"):
        a, b = ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]
        # a coarse Coulomb grid interpolates far below zero next to capped
        # nodes (no overlap short-cut, as in the reference), so some W
        # overflow: the same ones on both sides
        fin = np.isfinite(a)
        assert np.array_equal(fin, np.isfinite(b)), f
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[np.isinf(a)], b[np.isinf(b)])
        a, b = a[fin], b[fin]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) <= 1e-9, f


def test_grid_widom_tracks_direct(W, O, R):
    """At RASPA2's default spacing (0.15 A) the Widom weight from grids stays
    within 1% of the direct sums over the same trials (the reference's own
    grid check warns above 1%, src/grids.c:1640-1660)."""
    fw = W.c1_framework()
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    tg, mg = W.register_probe(ctx, W.CO2)
    ctx0 = fw.context(charge_method=R.CHARGE_EWALD)
    _, mg0 = W.register_probe(ctx0, W.CO2)
    go = ctx.MakeGrid(tg[0], fw.unit_box, 0.15)
    gcb = ctx.MakeGrid(tg[1], fw.unit_box, 0.15)
    gc = ctx.MakeCoulombGrid(fw.unit_box, 0.15)
    for t, g in ((tg[0], go), (tg[1], gcb), (tg[2], go)):
        ctx.use_grid(t, g)
    ctx.use_coulomb_grid(gc)
    _, rec = ctx.widom_run(mg, 5, 50000, records=True)
    _, rec0 = ctx0.widom_run(mg0, 5, 50000, records=True)
    assert np.all(np.isfinite(rec[:, 0]))
    assert abs(rec[:, 0].mean() / rec0[:, 0].mean() - 1.0) < 0.01
