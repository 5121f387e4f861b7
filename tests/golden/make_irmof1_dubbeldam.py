"""Writes tests/golden/irmof1_dubbeldam.json: the parameters of the
reference's grid example decks (examples/Auxiliary/MakingGrids and
UsingGrids: IRMOF-1, force field Dubbeldam2007FlexibleIRMOF-1, grids for
C_co2 and O_co2 plus the Ewald-Coulomb grid, CO2 from TraPPE).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_irmof1_dubbeldam.py

Inputs (read as data, nothing is executed):
  examples/Auxiliary/MakingGrids/simulation.input, UsingGrids/simulation.input
  forcefield/Dubbeldam2007FlexibleIRMOF-1/force_field_mixing_rules.def  LJ eps/sigma
  forcefield/Dubbeldam2007FlexibleIRMOF-1/pseudo_atoms.def              charges, masses
  molecules/TraPPE/CO2.def                                              rigid CO2 geometry
The atom positions are those of tests/golden/irmof1.json (the P1 expansion of
structures/mofs/cif/IRMOF-1.cif), whose labels name these pseudo-atoms.
"""
import json
import os

REF = "/root/reference"
FF = os.path.join(REF, "forcefield/Dubbeldam2007FlexibleIRMOF-1")
DECKS = os.path.join(REF, "examples/Auxiliary")
HERE = os.path.dirname(os.path.abspath(__file__))


def deck(path):
    """simulation.input as {keyword: [values]} (first occurrence)."""
    out = {}
    for ln in open(path):
        p = ln.split()
        if p and p[0] not in out:
            out[p[0]] = p[1:]
    return out


def mixing_rules():
    lines = [ln.split() for ln in open(os.path.join(FF, "force_field_mixing_rules.def"))]
    body = [p for p in lines if p and not p[0].startswith("#")]
    shifted = body[0][0] == "shifted"
    lj = {p[0]: (float(p[2]), float(p[3])) for p in body if len(p) >= 4 and p[1] == "LENNARD_JONES"}
    return shifted, lj


def pseudo_atoms():
    rows = {}
    for ln in open(os.path.join(FF, "pseudo_atoms.def")):
        p = ln.split()
        if len(p) >= 7 and not p[0].startswith("#"):
            try:
                rows.setdefault(p[0], {"mass": float(p[5]), "charge": float(p[6])})
            except ValueError:
                pass
    return rows


def co2():
    lines = [ln.split() for ln in open(os.path.join(REF, "molecules/TraPPE/CO2.def"))]
    sites = [p for p in lines if len(p) == 5 and p[0].isdigit() and p[1].endswith("_co2")]
    return [(p[1], [float(p[2]), float(p[3]), float(p[4])]) for p in sites]


def main():
    making = deck(os.path.join(DECKS, "MakingGrids/simulation.input"))
    using = deck(os.path.join(DECKS, "UsingGrids/simulation.input"))
    shifted, lj = mixing_rules()
    pa = pseudo_atoms()
    fw = json.load(open(os.path.join(HERE, "irmof1.json")))
    labels = sorted({a["label"] for a in fw["atoms"]})
    out = {
        "source": {"decks": ["examples/Auxiliary/MakingGrids/simulation.input",
                             "examples/Auxiliary/UsingGrids/simulation.input"],
                   "forcefield": "forcefield/Dubbeldam2007FlexibleIRMOF-1",
                   "molecule": "molecules/TraPPE/CO2.def",
                   "positions": "tests/golden/irmof1.json"},
        "forcefield": making["Forcefield"][0],
        "framework": making["FrameworkName"][0],
        "unit_cells": [int(v) for v in making["UnitCells"]],
        "grid_types": making["GridTypes"],
        "spacing_vdw": float(making["SpacingVDWGrid"][0]),
        "spacing_coulomb": float(making["SpacingCoulombGrid"][0]),
        "charge_method": using["ChargeMethod"][0],
        "ewald_precision": float(using["EwaldPrecision"][0]),
        "temperature": float(using["ExternalTemperature"][0]),
        "shifted": shifted,
        "framework_types": {l: {"eps_k": lj[l][0], "sigma": lj[l][1], **pa[l]} for l in labels},
        "probe_types": {t: {"eps_k": lj[t][0], "sigma": lj[t][1], **pa[t]}
                        for t in making["GridTypes"]},
        "co2": co2(),
    }
    with open(os.path.join(HERE, "irmof1_dubbeldam.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("forcefield", "grid_types", "spacing_vdw", "framework_types")}))


if __name__ == "__main__":
    main()
