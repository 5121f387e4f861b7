"""Pins of the oracle against numbers the reference itself holds: the
printed unit factors, mixed LJ parameters, shifts, cell volume, framework
mass/density and Widom estimators of the He/IRMOF-1 example output
(examples/Auxiliary/HeliumVoidFraction/IRMOF-1/Output/System_0/
output_IRMOF-1_1.1.1_298.000000_0.data, lines cited in tests/golden/irmof1.json),
and the C++ standard's known answer for MT19937-64."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "irmof1.json")))
OUT = G["output"]


def irmof1_system(O, T=298.0):
    a, b, c, al, be, ga = G["cell"]
    box, inv, props = O.cell_make(a, b, c, al, be, ga)
    frac = np.array([at["frac"] for at in G["atoms"]])
    xyz = frac @ box.reshape(3, 3)
    n = len(frac)
    s = O.System(box, O.BC_TRICLINIC, xyz[:, 0], xyz[:, 1], xyz[:, 2], np.zeros(n),
                 [at["eps_k"] for at in G["atoms"]], [at["sigma"] for at in G["atoms"]],
                 mass=[at["mass"] for at in G["atoms"]], cutoff_vdw=G["cutoff"], shifted=True,
                 temperature=T)
    return s, box, props


def test_units_printed(O):
    u = O.units()
    assert round(u["energy_to_kelvin"], 10) == OUT["energy_to_kelvin"]
    assert round(u["coulomb_kelvin"], 10) == OUT["coulomb_potential_unit_k"]
    assert round(u["k_b"], 10) == OUT["boltzmann_internal"]


def test_mt19937_64_known_answer(O):
    # 10000th output of the default-seeded (5489) generator, C++ [rand.predef]
    assert int(O.mt64(5489, 10000)[-1]) == 9981545732273789042


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_mt19937_64_against_std(O):
    """init_genrand64 seeding and the output function equal std::mt19937_64."""
    src = r'''#include <random>
#include <cstdio>
#include <cstdlib>
int main(int c, char** v){ std::mt19937_64 g(strtoull(v[1],0,10));
 for(int i=0;i<3000;i++) printf("%llu\n",(unsigned long long)g()); }'''
    with tempfile.TemporaryDirectory() as d:
        cpp, exe = os.path.join(d, "m.cpp"), os.path.join(d, "m")
        open(cpp, "w").write(src)
        subprocess.check_call(["g++", "-O1", "-o", exe, cpp])
        for seed in (0, 1, 1417341545, 2 ** 63 + 12345):
            out = subprocess.check_output([exe, str(seed)]).decode().split()
            ref = np.array([int(x) for x in out], dtype=np.uint64)
            assert np.array_equal(O.mt64(seed, 3000), ref), seed


def test_random_number_closed_interval(O):
    u = O.mt64_uniform(7, 1000)
    x = O.mt64(7, 1000)
    assert np.array_equal(u, (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740991.0))


def test_mixing_and_shift_printed(O):
    """Lorentz-Berthelot mixing and the cutoff shift, printed to 5 and 8
    decimals by the reference for He against each framework type."""
    s, box, props = irmof1_system(O)
    e2k = O.units()["energy_to_kelvin"]
    he = s.add_pseudo_atom(G["helium"]["eps_k"], G["helium"]["sigma"])
    label_atom = {}
    for i, at in enumerate(G["atoms"]):
        label_atom.setdefault(at["label"], i)
    for key, ref in OUT["pairs"].items():
        b = key.split("-")[1]
        if b == "He":
            continue
        eps, sig, shift = s.pair_params(he, label_atom[b])
        assert round(eps * e2k, 5) == ref["eps_k"], key
        assert round(sig, 5) == round(ref["sigma"], 5) or abs(sig - ref["sigma"]) <= 5e-6, key
        assert round(shift * e2k, 8) == ref["shift_k"], key


def test_cell_volume_mass_density_printed(O):
    s, box, props = irmof1_system(O)
    assert round(props[6], 9) == round(OUT["volume"], 9)
    mass = sum(at["mass"] for at in G["atoms"])
    # printed masses carry 9 decimals, so the sum agrees to ~1e-7 g/mol
    assert abs(mass - OUT["framework_mass"]) < 1e-6
    rho = O.framework_density(OUT["framework_mass"], props[6])
    assert round(rho, 9) == round(OUT["framework_density"], 9)


def test_henry_formula_printed(O):
    """K_H from <W> with the reference's own estimator: 0.803668 -> 5.46615e-07
    mol/kg/Pa at 298 K (printed to 6 digits)."""
    blocks = np.zeros((5, 4))
    for k, (w, _) in enumerate(OUT["widom_w_blocks"]):
        blocks[k] = [1.0, w, 0.0, 1.0]
    est = O.widom_estimators(blocks, 298.0, OUT["framework_density"])
    assert float("%.6g" % est["kh"]) == OUT["henry"]


def test_helium_widom_statistical(O):
    """He Widom in IRMOF-1 with 10 first-bead trials, 2e4 insertions of the
    counter stream: <W> agrees with the reference's 1e7-insertion result
    0.803668 +/- 0.000255 within 5 combined standard errors."""
    s, box, props = irmof1_system(O)
    he = s.add_pseudo_atom(G["helium"]["eps_k"], G["helium"]["sigma"])
    mol = O.molecule([he], [[0, 0, 0]], start=0, k_fb=10, k_or=10)
    rec = O.widom_run_stream(s, mol, 2024, 0, 20000)
    w = rec[:, 0]
    se = w.std(ddof=1) / np.sqrt(len(w))
    assert abs(w.mean() - OUT["widom_w"]) < 5 * np.hypot(se, OUT["widom_w_2sigma"] / 2)
