"""Writes tests/golden/ideal_rosenbluth_c5_c9.json: the reference's example
examples/Auxiliary/IdealRosenbluthWeights-C5-C9-573K (TraPPE n-pentane to
n-nonane grown by CBMC in an empty 30 A box at 573 K, force field
GarciaPerez2006, 25000 cycles, Widom probability 1 for each component) and
the average Widom Rosenbluth weights it printed, and copies the data files
the test reads into tests/golden/share/raspa.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_ideal_rosenbluth.py

Inputs (read as data, nothing is executed):
  examples/Auxiliary/IdealRosenbluthWeights-C5-C9-573K/simulation.input
  examples/Auxiliary/IdealRosenbluthWeights-C5-C9-573K/Output/System_0/output_Box_1.1.1_573.000000_0.data
  forcefield/GarciaPerez2006/{force_field_mixing_rules,pseudo_atoms,force_field}.def
  molecules/TraPPE/C5.def ... C9.def
"""
import json
import os
import re
import shutil

REF = "/root/reference"
EX = os.path.join(REF, "examples/Auxiliary/IdealRosenbluthWeights-C5-C9-573K")
OUT = os.path.join(EX, "Output/System_0/output_Box_1.1.1_573.000000_0.data")
HERE = os.path.dirname(os.path.abspath(__file__))
SHARE = os.path.join(HERE, "share", "raspa")


def main():
    deck = {}
    for ln in open(os.path.join(EX, "simulation.input")):
        p = ln.split()
        if p and p[0] not in deck:
            deck[p[0]] = p[1:]
    lines = open(OUT).read().splitlines()
    settings = {}
    for i, ln in enumerate(lines):
        m = re.match(r"\s*Number of (trial positions \(Widom insertion\)|trial positions coupled Torsion-selection"
                     r"|trial positions first bead \(Widom insertion\)|trial moves per open bead):\s+(\d+)", ln)
        if m:
            settings[m.group(1)] = {"value": int(m.group(2)), "line": i + 1}
        m = re.match(r"CutOff VDW : ([0-9.]+)", ln)
        if m:
            settings["CutOff VDW"] = {"value": float(m.group(1)), "line": i + 1}
    weights = {}
    for i, ln in enumerate(lines):
        m = re.match(r"\s*\[(C\d)\] Average Widom:\s+([0-9.eE+-]+) \+/- ([0-9.eE+-]+)", ln)
        if m:
            blocks = [float(lines[j].split()[2]) for j in range(i - 6, i - 1)]
            weights[m.group(1)] = {"average": float(m.group(2)), "plus_minus": float(m.group(3)),
                                   "blocks": blocks, "line": i + 1}
    # the run's own bonded parameters, as printed per component (the
    # molecule files shipped today carry p_1 = 355.03 K for TRAPPE_DIHEDRAL,
    # the run printed 335.03 K)
    torsions, comp = {}, None
    for i, ln in enumerate(lines):
        m = re.match(r"Component \d+ \[(C\d)\]", ln)
        if m:
            comp = m.group(1)
        m = re.search(r"TRAPPE_DIHEDRAL: p_0/k_B=([0-9.eE+-]+)\s+\[K\], p_1/k_B=([0-9.eE+-]+) \[K\], "
                      r"p_2/k_B=([0-9.eE+-]+) \[K\], p_3/k_B=([0-9.eE+-]+)", ln)
        if m and comp and comp not in torsions:
            torsions[comp] = {"p": [float(m.group(k)) for k in range(1, 5)], "line": i + 1}
    lj = {}
    for i, ln in enumerate(lines):
        m = re.match(r"(\S+) - (\S+) \[LENNARD_JONES\] p_0/k_B:\s+([0-9.]+) \[K\], p_1: ([0-9.]+) \[A\]", ln)
        if m and m.group(1) in ("CH3_sp3", "CH2_sp3") and m.group(2) in ("CH3_sp3", "CH2_sp3"):
            lj[m.group(1) + "-" + m.group(2)] = {"eps_k": float(m.group(3)), "sigma": float(m.group(4)),
                                                 "line": i + 1}
    doc = {
        "source": "examples/Auxiliary/IdealRosenbluthWeights-C5-C9-573K",
        "output": os.path.relpath(OUT, REF),
        "temperature": float(deck["ExternalTemperature"][0]),
        "box": [float(v) for v in deck["BoxLengths"]],
        "forcefield": deck["Forcefield"][0],
        "cycles": int(deck["NumberOfCycles"][0]),
        "settings": settings,
        "ideal_chain_rosenbluth_weight_printed": 1.0,
        "average_widom": weights,
        "printed_torsions": torsions,
        "printed_lj": lj,
    }
    with open(os.path.join(HERE, "ideal_rosenbluth_c5_c9.json"), "w") as f:
        json.dump(doc, f, indent=1)
    ff = os.path.join(SHARE, "forcefield", "GarciaPerez2006")
    os.makedirs(ff, exist_ok=True)
    for name in ("force_field_mixing_rules.def", "pseudo_atoms.def", "force_field.def"):
        shutil.copy(os.path.join(REF, "forcefield/GarciaPerez2006", name), ff)
    mol = os.path.join(SHARE, "molecules", "TraPPE")
    os.makedirs(mol, exist_ok=True)
    for n in range(4, 10):
        shutil.copy(os.path.join(REF, f"molecules/TraPPE/C{n}.def"), mol)


if __name__ == "__main__":
    main()
