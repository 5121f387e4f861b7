"""The reference's helium void-fraction example end to end on the GPU:
examples/Auxiliary/HeliumVoidFraction (IRMOF-1 from its CIF, GenericMOFs,
TraPPE helium, 298 K, 12 A cut-off) read through the front-end, against the
reference's printed 1e7-insertion results (output lines cited in
tests/golden/irmof1.json), and the wrapper call of python/raspa2.py:345."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
OUT = json.load(open(os.path.join(GOLD, "irmof1.json")))["output"]


def _cif():
    return open(os.path.join(GOLD, "share", "raspa", "structures", "cif", "IRMOF-1.cif")).read()


def test_helium_void_fraction_example(R):
    """<W>, K_H and <U_gh> of 4e6 helium insertions (10 first-bead trials
    each) agree with the reference's printed averages within three times the
    two printed 2-sigma spreads combined."""
    from raspa2_amd import frontend as F
    est = F.widom(_cif(), "helium", forcefield="GenericMOFs", cutoff=12.0, temperature=298.0,
                  raspa_dir=GOLD, insertions=4_000_000, seed=11, precision=R.PREC_FP64)
    assert abs(est["w"] - OUT["widom_w"]) < 3 * (OUT["widom_w_2sigma"] + est["w_err"])
    assert abs(est["kh"] - OUT["henry"]) < 3 * (OUT["henry_2sigma"] + est["kh_err"]) + 1e-3 * OUT["henry"]
    assert abs(est["u"] - OUT["u_gh_k"]) < 3 * (OUT["u_gh_2sigma_k"] + est["u_err"])


def test_wrapper_call(R):
    """get_helium_void_fraction with the wrapper's own settings (CrystalGenerator
    force field, 12.8 A cut-off, 2000 cycles = 40000 insertions)."""
    from raspa2_amd import frontend as F
    w = F.get_helium_void_fraction(_cif(), cycles=2000, forcefield="CrystalGenerator", raspa_dir=GOLD)
    assert 0.7 < w < 0.9


def test_surface_area_device_equals_oracle(W, O, R):
    """SurfaceAreaMove on the device: the accessible points of every atom
    equal the oracle's on the same counter-stream points (IRMOF-1 with the
    N2 probe atom, RASPA2's default distance 2^(1/6) sigma and 'Sigma')."""
    from raspa2_amd import frontend as F
    share = os.path.join(GOLD, "share", "raspa")
    ff = F.ForceField.load(os.path.join(share, "forcefield", "GenericMOFs"))
    fw = F.framework_from_cif(_cif(), ff)
    n2 = F.read_molecule(os.path.join(share, "molecules", "TraPPE", "N2.def"), ff)
    ctx = fw.context()
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    s = O.System(box, bc, x, y, z, q, eps, sig, mass=mass)
    tg = ctx.add_pseudo_atom(*n2.sites[0])
    to = s.add_pseudo_atom(*n2.sites[0])
    for factor in (2.0 ** (1.0 / 6.0), 1.0):
        ag, cg = ctx.SurfaceArea(tg, 300, seed=7, distance_factor=factor)
        ao, co = s.surface_area(to, 300, seed=7, factor=factor)
        assert np.array_equal(cg, co)
        assert abs(ag - ao) <= 1e-12 * ao
        assert 0 < cg.sum() < 300 * fw.n
    # atoms of NONE pseudo-atoms (here every hydrogen) carry no sphere and
    # exclude no point
    mask = np.array([not lab.startswith("H") for lab in fw.labels], dtype=np.uint8)
    assert 0 < mask.sum() < fw.n
    ag, cg = ctx.SurfaceArea(tg, 300, seed=7, distance_factor=1.0, interacts=mask)
    ao, co = s.surface_area(to, 300, seed=7, factor=1.0, interacts=mask)
    assert np.array_equal(cg, co) and not cg[mask == 0].any()
    assert abs(ag - ao) <= 1e-12 * ao


def test_surface_area_wrapper_call(R):
    """get_geometric_surface_area with the wrapper's settings: IRMOF-1 in
    m^2/g (the wrapper's default units) in the range reported for it."""
    from raspa2_amd import frontend as F
    sa = F.get_geometric_surface_area(_cif(), cycles=500, raspa_dir=GOLD)
    assert 3000.0 < sa < 4200.0, sa
    a2 = F.get_geometric_surface_area(_cif(), cycles=500, units="A^2", raspa_dir=GOLD)
    v = F.get_geometric_surface_area(_cif(), cycles=500, units="m^2/cm^3", raspa_dir=GOLD)
    assert abs(v - a2 / 25.832 ** 3 * 1e4) <= 1e-9 * v


def test_flexible_chain_widom_wrapper(R, W):
    """n-butane (molecules/TraPPE/C4.def) in IRMOF-1 through the wrapper
    call: the blocks equal a direct chain run with the same seed, and the
    Henry coefficient divides by the ideal-gas Rosenbluth weight, which is the
    empty-box mean of exp(-beta U_torsion) for butane within its error."""
    import math
    from raspa2_amd import frontend as F
    n = 200_000
    w_ig = F.ideal_gas_rosenbluth_weight("C4", forcefield="GenericMOFs", temperature=298.0, cutoff=12.0,
                                         raspa_dir=GOLD, insertions=400_000)
    phi = np.linspace(-math.pi, math.pi, 400001)
    u = 355.03 * (1 + np.cos(phi)) - 68.19 * (1 - np.cos(2 * phi)) + 791.32 * (1 + np.cos(3 * phi))
    # RosenbluthTorsion averages 10 of the 100 torsion trials: same mean
    assert w_ig == pytest.approx(np.exp(-u / 298.0).mean(), rel=0.01)
    est = F.widom(_cif(), "C4", forcefield="GenericMOFs", cutoff=12.0, temperature=298.0, raspa_dir=GOLD,
                  insertions=n, seed=3, precision=R.PREC_FP64, ideal_gas_weight=w_ig)
    est1 = F.widom(_cif(), "C4", forcefield="GenericMOFs", cutoff=12.0, temperature=298.0, raspa_dir=GOLD,
                   insertions=n, seed=3, precision=R.PREC_FP64)
    assert est["kh_overall"] == pytest.approx(est1["kh_overall"] / w_ig, rel=1e-12)
    assert est["w"] > w_ig  # butane is attracted into the framework
    assert est["u"] < 0.0
