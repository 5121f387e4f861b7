"""The reference's grid example decks on the GPU: MakingGrids (IRMOF-1,
Dubbeldam2007FlexibleIRMOF-1, grids for C_co2 and O_co2 and the Ewald-Coulomb
grid, written where RASPA2 keeps them) and UsingGrids (the files read back,
CO2 Widom insertion served from them), against the oracle on the same grids
(examples/Auxiliary/MakingGrids, UsingGrids; src/grids.c:431-975)."""
import os

import numpy as np
import pytest

from helpers import irmof1_dubbeldam, oracle_molecule, oracle_system, gpu_molecule, random_positions

pytestmark = pytest.mark.gpu

SPACING = 0.5  # the decks use 0.1; the oracle checks sample a coarser grid
FIELDS = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
              FBMASK=9, ORMASK=10, STAGE=13)


def _context(fw, D, R):
    return fw.context(charge_method=R.CHARGE_EWALD, ewald_precision=D["ewald_precision"],
                      temperature=D["temperature"])


def _make(fw, probe, D, R, root):
    """MakeGrid as the deck runs it; returns the paths written."""
    ctx = _context(fw, D, R)
    names = [n for n, _ in D["co2"]]
    types = {n: ctx.add_pseudo_atom(*probe.sites[names.index(n)]) for n in D["grid_types"]}
    paths = {}
    for n in D["grid_types"]:
        g = ctx.MakeGrid(types[n], fw.unit_box, SPACING)
        p = R.grid_path(root, D["forcefield"], D["framework"], SPACING, pseudo_atom=n, shifted=D["shifted"])
        os.makedirs(os.path.dirname(p), exist_ok=True)
        g.write(p)
        paths[n] = (p, g.download())
        g.close()
    g = ctx.MakeCoulombGrid(fw.unit_box, SPACING)
    p = R.grid_path(root, D["forcefield"], D["framework"], SPACING, n_unit_cells=D["unit_cells"])
    os.makedirs(os.path.dirname(p), exist_ok=True)
    g.write(p)
    paths["coulomb"] = (p, g.download())
    g.close()
    ctx.close()
    return paths


def test_making_grids_deck(W, O, R, tmp_path):
    """The deck's grids land under RASPA2's file names; sampled nodes equal
    the oracle's (VDW bit for bit, Coulomb within two float ulps)."""
    fw, probe, D = irmof1_dubbeldam()
    assert abs(fw.charge.sum()) < 1e-9
    paths = _make(fw, probe, D, R, str(tmp_path))
    base = os.path.join(str(tmp_path), "share", "raspa", "grids", D["forcefield"], "IRMOF-1", "0.500000")
    assert paths["C_co2"][0] == os.path.join(base, "IRMOF-1_C_co2_shifted.grid")
    assert paths["coulomb"][0] == os.path.join(base, "1x1x1", "IRMOF-1_Electrostatics_Ewald.grid")
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD, precision=D["ewald_precision"],
                          T=D["temperature"])
    names = [n for n, _ in D["co2"]]
    geom = O.grid_geometry(fw.unit_box, SPACING)
    rng = np.random.default_rng(71)
    for n in D["grid_types"]:
        t = sys_o.add_pseudo_atom(*probe.sites[names.index(n)])
        data = paths[n][1]
        for _ in range(80):
            i, j, k = (int(rng.integers(0, s)) for s in data.shape[:3])
            ref = sys_o.grid_node(t, geom, i, j, k)
            assert np.array_equal(data[i, j, k].view(np.uint32), ref.view(np.uint32)), (n, i, j, k)
    data = paths["coulomb"][1]
    for _ in range(80):
        i, j, k = (int(rng.integers(0, s)) for s in data.shape[:3])
        ref = sys_o.coulomb_grid_node(geom, i, j, k)
        got = data[i, j, k]
        assert (got[0] == np.float32(2e7)) == (ref[0] == np.float32(2e7))
        tol = 2.0 * np.spacing(np.abs(ref)) + 1e-9 * np.abs(ref).max()
        assert np.all(np.abs(got.astype(np.float64) - ref) <= tol), (i, j, k)


def test_using_grids_deck(W, O, R, tmp_path):
    """UsingGrids: the files read onto a fresh context serve CO2 Widom
    insertion; records equal the oracle's on the same grid data."""
    fw, probe, D = irmof1_dubbeldam()
    paths = _make(fw, probe, D, R, str(tmp_path))
    ctx = _context(fw, D, R)
    tg, mg = gpu_molecule(ctx, probe)
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD, precision=D["ewald_precision"],
                          T=D["temperature"])
    to, mo = oracle_molecule(sys_o, probe)
    geom = O.grid_geometry(fw.unit_box, SPACING)
    names = [n for n, _ in D["co2"]]
    keep = []
    for n in D["grid_types"]:
        g = ctx.ReadGrid(paths[n][0], fw.unit_box)
        keep.append(g)
        assert g.spacing == SPACING and np.array_equal(g.download().view(np.uint32),
                                                       paths[n][1].view(np.uint32))
        for j, name in enumerate(names):
            if name == n:
                ctx.use_grid(tg[j], g)
                sys_o.use_grid(to[j], paths[n][1], geom, fw.unit_box)
    gc = ctx.ReadGrid(paths["coulomb"][0], fw.unit_box, coulomb=True, n_unit_cells=D["unit_cells"])
    assert gc.ewald_precision == D["ewald_precision"]
    ctx.use_coulomb_grid(gc)
    sys_o.use_coulomb_grid(paths["coulomb"][1], geom, fw.unit_box)
    ro = O.widom_run_stream(sys_o, mo, 23, 0, 300)
    _, rg = ctx.widom_run(mg, 23, 300, records=True)
    for f in ("OVERLAP", "FBMASK", "ORMASK", "STAGE", "SELFB", "SELOR"):
        assert np.array_equal(ro[:, FIELDS[f]], rg[:, FIELDS[f]]), f
    acc = ro[:, FIELDS["STAGE"]] == 4
    assert acc.sum() > 50
    for f in ("UVDW", "UREAL", "UFHOST", "UFADS", "DU", "W"):
        a, b = ro[acc, FIELDS[f]], rg[acc, FIELDS[f]]
        fin = np.isfinite(a)
        assert np.array_equal(fin, np.isfinite(b)), f
        a, b = a[fin], b[fin]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) <= 1e-9, f
