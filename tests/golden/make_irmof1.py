"""Writes tests/golden/irmof1.json from the reference's own input and output
files for the He-in-IRMOF-1 Widom example.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_irmof1.py

Inputs (read as data, nothing is executed):
  structures/mofs/cif/IRMOF-1.cif                 cell, symmetry operators, atom sites
  forcefield/GenericMOFs/force_field_mixing_rules.def   LJ parameters (Zn_, O_, C_, H_, He)
  examples/Auxiliary/HeliumVoidFraction/IRMOF-1/Output/System_0/
      output_IRMOF-1_1.1.1_298.000000_0.data      framework masses, printed mixed LJ
                                                   parameters and shifts, Widom results
The JSON holds the P1 expansion of the CIF (424 atoms) and the printed
numbers the tests compare against, each with its output-file line.
"""
import json
import os
import re
from fractions import Fraction

REF = "/root/reference"
CIF = os.path.join(REF, "structures/mofs/cif/IRMOF-1.cif")
FF = os.path.join(REF, "forcefield/GenericMOFs/force_field_mixing_rules.def")
OUT = os.path.join(REF, "examples/Auxiliary/HeliumVoidFraction/IRMOF-1/Output/System_0/"
                        "output_IRMOF-1_1.1.1_298.000000_0.data")
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_op(op):
    """'x+1/2,-y,z' -> list of (coef_x, coef_y, coef_z, const) per component."""
    rows = []
    for comp in op.replace(" ", "").split(","):
        coef = {"x": 0, "y": 0, "z": 0}
        const = Fraction(0)
        for sign, term in re.findall(r"([+-]?)([^+-]+)", comp):
            s = -1 if sign == "-" else 1
            if term in coef:
                coef[term] += s
            else:
                const += s * Fraction(term)
        rows.append((coef["x"], coef["y"], coef["z"], float(const)))
    return rows


def read_cif():
    lines = open(CIF).read().splitlines()
    cell = {}
    for ln in lines:
        for key in ("_cell_length_a", "_cell_length_b", "_cell_length_c", "_cell_angle_alpha",
                    "_cell_angle_beta", "_cell_angle_gamma"):
            if ln.startswith(key + " ") or ln.startswith(key + "\t"):
                cell[key] = float(ln.split()[1])
    ops = [ln.strip().strip("'") for ln in lines if re.match(r"^\s*'[^']*,[^']*,[^']*'\s*$", ln)]
    sites = []
    in_sites = False
    for ln in lines:
        if ln.startswith("_atom_site_charge"):
            in_sites = True
            continue
        if in_sites:
            p = ln.split()
            if len(p) >= 5:
                sites.append((p[0], p[1], float(p[2]), float(p[3]), float(p[4])))
    return cell, [parse_op(o) for o in ops], sites


def expand(ops, sites):
    atoms = []
    for label, elem, fx, fy, fz in sites:
        for op in ops:
            f = []
            for cx, cy, cz, c0 in op:
                v = cx * fx + cy * fy + cz * fz + c0
                f.append(v - float(int(v // 1)))
            f = [v - 1.0 if v >= 1.0 else v for v in f]
            dup = False
            for a in atoms:
                if a[0] != label:
                    continue
                d = [abs(f[k] - a[2][k]) for k in range(3)]
                d = [min(x, 1.0 - x) for x in d]
                if max(d) < 1e-4:
                    dup = True
                    break
            if not dup:
                atoms.append((label, elem, f))
    return atoms


def read_ff():
    lj = {}
    for ln in open(FF).read().splitlines():
        p = ln.split()
        if len(p) >= 4 and p[1].lower() == "lennard-jones":
            lj[p[0]] = (float(p[2]), float(p[3]))
    return lj


def read_output():
    lines = open(OUT).read().splitlines()
    res = {"file": os.path.relpath(OUT, REF)}

    def grab(pattern, conv=float):
        for i, ln in enumerate(lines):
            m = re.search(pattern, ln)
            if m:
                return [conv(g) for g in m.groups()], i + 1
        raise RuntimeError(pattern)

    masses = {}
    for lab in ("Zn1", "O1", "O2", "C1", "C2", "C3", "H1"):
        (m,), ln = grab(r"Name %s\s.*Mass=([0-9.]+)" % lab)
        masses[lab] = {"mass": m, "line": ln}
    res["masses"] = masses
    pairs = {}
    for a, b in (("He", "Zn1"), ("He", "O1"), ("He", "C1"), ("He", "H1"), ("He", "He")):
        (e, s, sh), ln = grab(r"^\s*%s -\s+%s \[LENNARD_JONES\] p_0/k_B:\s+([-0-9.]+) \[K\], "
                              r"p_1:\s+([-0-9.]+) \[A\], shift/k_B:\s+([-0-9.]+)" % (a, b))
        pairs[f"{a}-{b}"] = {"eps_k": e, "sigma": s, "shift_k": sh, "line": ln}
    res["pairs"] = pairs
    (v,), res["volume_line"] = grab(r"^volume of the cell: ([0-9.]+)")
    res["volume"] = v
    (m,), res["mass_line"] = grab(r"Framework Mass:\s+([0-9.]+)")
    res["framework_mass"] = m
    (d,), res["density_line"] = grab(r"Framework Density:\s+([0-9.]+)")
    res["framework_density"] = d
    (e2k,), res["e2k_line"] = grab(r"Energy to Kelvin:\s+([0-9.]+)")
    res["energy_to_kelvin"] = e2k
    (cu,), res["coulomb_unit_line"] = grab(r"Unit of Coulomb potential:\s+([0-9.]+)")
    res["coulomb_potential_unit_k"] = cu
    (kb,), res["kb_line"] = grab(r"Boltzmann constant:\s+([0-9.]+)")
    res["boltzmann_internal"] = kb
    (w, we), res["w_line"] = grab(r"\[helium\] Average Widom:\s+([-0-9.eE+]+) "
                                  r"\+/- ([-0-9.eE+]+)")
    res["widom_w"], res["widom_w_2sigma"] = w, we
    (kh, khe), res["kh_line"] = grab(r"\[helium\] Average Henry coefficient:\s+([-0-9.eE+]+) "
                                     r"\+/- ([-0-9.eE+]+)")
    res["henry"], res["henry_2sigma"] = kh, khe
    (u, ue), res["u_line"] = grab(r"\[helium\] Average\s+<U_gh>_1-<U_h>_0:\s+([-0-9.]+) "
                                  r"\+/-\s+([-0-9.]+)")
    blocks_w = []
    for i, ln in enumerate(lines):
        m = re.match(r"^\s*Block\[ ?(\d)\] ([0-9.]+) \[-\]", ln)
        if m:
            blocks_w.append((float(m.group(2)), i + 1))
    res["widom_w_blocks"] = blocks_w[-5:]
    res["u_gh_k"], res["u_gh_2sigma_k"] = u, ue
    (seed,), res["seed_line"] = grab(r"Random number seed: ([0-9]+)", int)
    res["seed"] = seed
    return res


def main():
    cell, ops, sites = read_cif()
    atoms = expand(ops, sites)
    lj = read_ff()
    out = read_output()
    prefix = {"Zn": "Zn_", "O": "O_", "C": "C_", "H": "H_"}
    data = {
        "source": {"cif": os.path.relpath(CIF, REF), "forcefield": os.path.relpath(FF, REF)},
        "cell": [cell["_cell_length_a"], cell["_cell_length_b"], cell["_cell_length_c"],
                 cell["_cell_angle_alpha"], cell["_cell_angle_beta"], cell["_cell_angle_gamma"]],
        "n_symmetry_ops": len(ops),
        "atoms": [{"label": lab, "elem": el, "frac": f, "eps_k": lj[prefix[el]][0],
                   "sigma": lj[prefix[el]][1], "mass": out["masses"][lab]["mass"]}
                  for lab, el, f in atoms],
        "helium": {"eps_k": lj["He"][0], "sigma": lj["He"][1]},
        "temperature": 298.0,
        "cutoff": 12.0,
        "output": out,
    }
    path = os.path.join(HERE, "irmof1.json")
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1)
    print(path, len(atoms), "atoms")


if __name__ == "__main__":
    main()
