"""Input front-end: RASPA2's structure, force-field and molecule files to the
library's framework and probe descriptions, and the reference Python
wrapper's convenience calls on top of the device Widom path.

Formats and rules followed:
  * CIF structures (src/framework.c:334-2050): cell lengths and angles,
    symmetry operators (`_symmetry_equiv_pos_as_xyz` or
    `_space_group_symop_operation_xyz`), atom sites (label, type symbol,
    fractional coordinates, charge); the pseudo-atom of a site is its label;
    a site without a type symbol takes the letters of its label.
  * force_field_mixing_rules.def (src/potentials.c:2990-3130): shifted or
    truncated, tail corrections, self parameters per pseudo-atom in file
    order, a name ending in '_' applying to every pseudo-atom whose name
    starts with the rest (case-insensitive), later lines overriding earlier
    ones; Lorentz-Berthelot mixing.
  * pseudo_atoms.def (src/potentials.c): mass and charge per pseudo-atom;
    these take precedence over the CIF file (src/framework.c:1575-1612).
  * molecule definitions (molecules/<set>/<name>.def): atom types and,
    for rigid groups, body coordinates.
  * boundary condition: rectangular when all cell angles are 90 degrees,
    triclinic otherwise.

This module parses text only; every energy is computed by librhg on the GPU.
"""
import os
import re
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from . import rhg as R
from . import workloads as W

# Standard atomic weights (IUPAC conventional values) for sites whose
# pseudo-atom pseudo_atoms.def does not define.
ATOMIC_MASS = {
    "H": 1.008, "He": 4.0026, "Li": 6.94, "Be": 9.0122, "B": 10.81, "C": 12.011, "N": 14.007,
    "O": 15.999, "F": 18.998, "Ne": 20.180, "Na": 22.990, "Mg": 24.305, "Al": 26.982, "Si": 28.085,
    "P": 30.974, "S": 32.06, "Cl": 35.45, "Ar": 39.948, "K": 39.098, "Ca": 40.078, "Sc": 44.956,
    "Ti": 47.867, "V": 50.942, "Cr": 51.996, "Mn": 54.938, "Fe": 55.845, "Co": 58.933, "Ni": 58.693,
    "Cu": 63.546, "Zn": 65.38, "Ga": 69.723, "Ge": 72.630, "As": 74.922, "Se": 78.971, "Br": 79.904,
    "Kr": 83.798, "Rb": 85.468, "Sr": 87.62, "Y": 88.906, "Zr": 91.224, "Nb": 92.906, "Mo": 95.95,
    "Ru": 101.07, "Rh": 102.91, "Pd": 106.42, "Ag": 107.87, "Cd": 112.41, "In": 114.82, "Sn": 118.71,
    "Sb": 121.76, "Te": 127.60, "I": 126.90, "Xe": 131.29, "Cs": 132.91, "Ba": 137.33, "La": 138.91,
    "Ce": 140.12, "Pr": 140.91, "Nd": 144.24, "Sm": 150.36, "Eu": 151.96, "Gd": 157.25, "Tb": 158.93,
    "Dy": 162.50, "Ho": 164.93, "Er": 167.26, "Tm": 168.93, "Yb": 173.05, "Lu": 174.97, "Hf": 178.49,
    "Ta": 180.95, "W": 183.84, "Re": 186.21, "Os": 190.23, "Ir": 192.22, "Pt": 195.08, "Au": 196.97,
    "Hg": 200.59, "Tl": 204.38, "Pb": 207.2, "Bi": 208.98, "U": 238.03,
}


# ------------------------------------------------------------------ CIF
def _tokens(text):
    """CIF tokens: quoted strings as one token, comments dropped."""
    out = []
    for line in text.splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        out.extend(m.group(1) if m.group(1) is not None else (m.group(2) if m.group(2) is not None else m.group(3))
                   for m in re.finditer(r"'([^']*)'|\"([^\"]*)\"|(\S+)", s))
    return out


def _number(tok):
    """CIF number with an optional standard uncertainty, e.g. 0.2934(3)."""
    return float(re.sub(r"\(\d+\)$", "", tok))


def parse_cif(text):
    """Cell, symmetry operators and atom sites of the first data block."""
    toks = _tokens(text)
    items, loops = {}, []
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.lower() == "loop_":
            tags = []
            i += 1
            while i < len(toks) and toks[i].startswith("_"):
                tags.append(toks[i].lower())
                i += 1
            vals = []
            while i < len(toks) and not toks[i].startswith("_") and toks[i].lower() != "loop_" \
                    and not toks[i].lower().startswith("data_"):
                vals.append(toks[i])
                i += 1
            rows = [vals[k:k + len(tags)] for k in range(0, len(vals) - len(tags) + 1, len(tags))]
            loops.append((tags, rows))
        elif t.startswith("_") and i + 1 < len(toks):
            items[t.lower()] = toks[i + 1]
            i += 2
        else:
            i += 1
    cell = tuple(_number(items["_cell_length_" + k]) for k in "abc") + \
        tuple(_number(items["_cell_angle_" + k]) for k in ("alpha", "beta", "gamma"))
    ops, sites = [], []
    for tags, rows in loops:
        for key in ("_symmetry_equiv_pos_as_xyz", "_space_group_symop_operation_xyz"):
            if key in tags:
                ops = [r[tags.index(key)] for r in rows]
        if "_atom_site_fract_x" in tags:
            col = {t: tags.index(t) for t in tags}
            for r in rows:
                label = r[col["_atom_site_label"]] if "_atom_site_label" in col else ""
                sym = r[col["_atom_site_type_symbol"]] if "_atom_site_type_symbol" in col else ""
                if not sym:
                    sym = re.match(r"[A-Za-z]*", label).group(0)
                if not label:
                    label = re.match(r"[A-Za-z]*", sym).group(0)
                q = None
                if "_atom_site_charge" in col:
                    try:
                        q = _number(r[col["_atom_site_charge"]])
                    except ValueError:
                        q = None
                sites.append((label, sym, _number(r[col["_atom_site_fract_x"]]),
                              _number(r[col["_atom_site_fract_y"]]), _number(r[col["_atom_site_fract_z"]]), q))
    if not ops:
        ops = ["x,y,z"]
    return {"cell": cell, "ops": ops, "sites": sites}


def _parse_op(op):
    """'x+1/2,-y,z' -> 3 rows of (coef_x, coef_y, coef_z, constant)."""
    rows = []
    for comp in op.replace(" ", "").lower().split(","):
        coef = {"x": 0, "y": 0, "z": 0}
        const = Fraction(0)
        for sign, term in re.findall(r"([+-]?)([^+-]+)", comp):
            s = -1 if sign == "-" else 1
            if term in coef:
                coef[term] += s
            elif term[-1] in coef:  # e.g. 2x
                coef[term[-1]] += s * int(term[:-1])
            else:
                const += s * Fraction(term)
        rows.append((coef["x"], coef["y"], coef["z"], float(const)))
    if len(rows) != 3:
        raise ValueError(f"bad symmetry operator {op!r}")
    return rows


def p1_atoms(cif, tol=1e-4):
    """Every image of every site under the symmetry operators, wrapped into
    [0, 1), duplicates (same label within tol in fractional coordinates,
    periodically) kept once, in site order then operator order."""
    ops = [_parse_op(o) for o in cif["ops"]]
    atoms = []
    for label, sym, fx, fy, fz, q in cif["sites"]:
        mine = []
        for op in ops:
            f = []
            for cx, cy, cz, c0 in op:
                v = cx * fx + cy * fy + cz * fz + c0
                v -= np.floor(v)
                f.append(0.0 if v >= 1.0 else v)
            dup = False
            for g in mine:
                d = np.abs(np.asarray(f) - g)
                if np.max(np.minimum(d, 1.0 - d)) < tol:
                    dup = True
                    break
            if not dup:
                mine.append(np.asarray(f))
        atoms.extend((label, sym, g, q) for g in mine)
    return atoms


# ------------------------------------------------------------ force field
def _data_lines(path):
    with open(path) as fh:
        return [ln.split() for ln in fh if ln.strip() and not ln.lstrip().startswith("#")]


def read_mixing_rules(path):
    """force_field_mixing_rules.def: (shifted, tail, [(name, potential,
    params)], mixing rule)."""
    rows = _data_lines(path)
    shifted = rows[0][0].lower() == "shifted"
    tail = rows[1][0].lower() == "yes"
    n = int(rows[2][0])
    def numbers(tok):
        # the parameters, up to a trailing comment (sscanf reads what it needs)
        out = []
        for t in tok:
            try:
                out.append(float(t))
            except ValueError:
                break
        return out

    entries = [(r[0], r[1].upper().replace("-", "_"), numbers(r[2:])) for r in rows[3:3 + n]]
    rule = rows[3 + n][0] if len(rows) > 3 + n else "Lorentz-Berthelot"
    return shifted, tail, entries, rule


def read_pseudo_atoms(path):
    """pseudo_atoms.def: {name: (element, mass, charge)}."""
    rows = _data_lines(path)
    n = int(rows[0][0])
    out = {}
    for r in rows[1:1 + n]:
        out.setdefault(r[0], (r[3], float(r[5]), float(r[6])))
    return out


@dataclass
class ForceField:
    """A RASPA2 force-field directory (share/raspa/forcefield/<name>)."""
    path: str
    shifted: bool = True
    tail: bool = False
    entries: list = field(default_factory=list)
    rule: str = "Lorentz-Berthelot"
    pseudo: dict = field(default_factory=dict)

    overrides: set = field(default_factory=set)

    def check_pairs(self, probe_labels, other_labels):
        """Refuses a calculation that forms a pair with an explicit
        force_field.def interaction (the kernels use Lorentz-Berthelot mixing
        of the self parameters)."""
        def hit(pat, name):
            if pat.endswith("_"):
                return name.lower().startswith(pat[:-1].lower())
            return name.lower() == pat.lower()

        others = set(other_labels) | set(probe_labels)
        for p1, p2 in self.overrides:
            for a in set(probe_labels):
                for b in others:
                    if (hit(p1, a) and hit(p2, b)) or (hit(p2, a) and hit(p1, b)):
                        raise NotImplementedError(f"force_field.def defines the pair {a}-{b}: explicit "
                                                  "pair interactions are not supported")

    @classmethod
    def load(cls, path):
        ff = cls(path)
        ff.shifted, ff.tail, ff.entries, ff.rule = read_mixing_rules(os.path.join(path, "force_field_mixing_rules.def"))
        pa = os.path.join(path, "pseudo_atoms.def")
        if os.path.exists(pa):
            ff.pseudo = read_pseudo_atoms(pa)
        over = os.path.join(path, "force_field.def")
        ff.overrides = set()
        if os.path.exists(over):
            # rules to overwrite, then explicitly defined pair interactions
            # (src/potentials.c:3000-3130); the kernels mix self parameters,
            # so a calculation that forms one of these pairs is refused
            rows = _data_lines(over)
            n_rules = int(rows[0][0])
            if n_rules:
                raise NotImplementedError("force_field.def mixing-rule overwrites are not supported")
            n_pairs = int(rows[1 + n_rules][0]) if len(rows) > 1 + n_rules else 0
            for r in rows[2 + n_rules:2 + n_rules + n_pairs]:
                ff.overrides.add(tuple(sorted((r[0], r[1]))))
        if ff.rule.lower() != "lorentz-berthelot":
            raise NotImplementedError(f"mixing rule {ff.rule}")
        if ff.tail:
            raise NotImplementedError("tail corrections")
        return ff

    def lj(self, name):
        """(eps/k_B [K], sigma [A]) of a pseudo-atom: every matching line in
        file order, exact names or '_' prefixes (src/potentials.c:3096-3118)."""
        got = None
        for pat, pot, par in self.entries:
            if pat.endswith("_"):
                hit = name.lower().startswith(pat[:-1].lower())
            else:
                hit = name.lower() == pat.lower()
            if hit:
                if pot == "NONE":
                    got = (0.0, 1.0)
                elif pot == "LENNARD_JONES":
                    got = (par[0], par[1])
                else:
                    raise NotImplementedError(f"potential {pot} for {name}")
        if got is None:
            raise KeyError(f"no force-field parameters for pseudo-atom {name!r}")
        return got

    def interacts(self, name):
        """False when the pseudo-atom's potential is NONE (PseudoAtoms[].Interaction,
        src/potentials.c:12170-12185), the last matching line deciding."""
        got = True
        for pat, pot, _ in self.entries:
            hit = name.lower().startswith(pat[:-1].lower()) if pat.endswith("_") else name.lower() == pat.lower()
            if hit:
                got = pot != "NONE"
        return got

    def mass_charge(self, name, element, cif_charge=None, use_cif_charges=False):
        if name in self.pseudo:
            _, m, q = self.pseudo[name]
            if use_cif_charges and cif_charge is not None:
                q = cif_charge
            return m, q
        el = re.match(r"[A-Za-z]*", element).group(0).capitalize()
        m = ATOMIC_MASS.get(el)
        if m is None:
            raise KeyError(f"no mass for element {element!r} of {name!r}")
        return m, (cif_charge if cif_charge is not None else 0.0)


def framework_from_cif(text, ff, unit_cells=(1, 1, 1), name="framework", use_cif_charges=False):
    """FrameworkSpec of unit_cells[0] x [1] x [2] cells of a CIF structure:
    the P1 atoms of the unit cell, replicated cell by cell."""
    cif = parse_cif(text)
    a, b, c, al, be, ga = cif["cell"]
    ubox = W.cell_box(a, b, c, al, be, ga)
    atoms = p1_atoms(cif)
    n = [int(v) for v in unit_cells]
    box = W.cell_box(a * n[0], b * n[1], c * n[2], al, be, ga)
    m = ubox.reshape(3, 3)
    frac = np.array([g for _, _, g, _ in atoms])
    ucart = frac @ m
    params = []
    for label, sym, _, q in atoms:
        eps, sig = ff.lj(label)
        mass, charge = ff.mass_charge(label, sym, q, use_cif_charges)
        params.append((eps, sig, charge, mass))
    params = np.array(params)
    xyz, P = [], []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                xyz.append(ucart + i * m[0] + j * m[1] + k * m[2])
                P.append(params)
    xyz = np.concatenate(xyz)
    P = np.concatenate(P)
    bc = R.BC_RECTANGULAR if (al, be, ga) == (90.0, 90.0, 90.0) else R.BC_TRICLINIC
    spec = W.FrameworkSpec(name, box, bc, xyz, P[:, 0], P[:, 1], P[:, 2], P[:, 3], unit_box=ubox)
    spec.labels = [lab for _ in range(int(np.prod(n))) for lab, _, _, _ in atoms]
    return spec


_BONDS = {"FIXED_BOND": R.BOND_FIXED, "HARMONIC_BOND": R.BOND_HARMONIC}
_BENDS = {"HARMONIC_BEND": R.BEND_HARMONIC}
_TORSIONS = {"TRAPPE_DIHEDRAL": R.TORSION_TRAPPE}


def read_molecule(path, ff, k_fb=10, k_or=10, start=0, k_torsion=100, moves_per_open_bead=150):
    """A molecule definition (ReadComponentDefinition, src/molecule.c:1430-3030).
    A rigid molecule (or a single atom) becomes a rigid probe with its body
    coordinates; a molecule made of one flexible group becomes a chain probe
    with its bonds, bends, torsions and intra-molecular LJ pairs, grown by
    CBMC with the internal MC scheme (k_or trial positions per bead).
    Pseudo-atom parameters come from the force field."""
    rows = []
    with open(path) as fh:
        rows = [ln.split() for ln in fh]
    data = [r for r in rows if r and not r[0].startswith("#")]
    n_atoms = int(data[3][0])
    n_groups = int(data[4][0])
    kind = data[5][0].lower()
    first = 7
    atoms = data[first:first + n_atoms]
    if len(atoms) != n_atoms or any(not (r[0].isdigit() and not r[1][0].isdigit()) for r in atoms):
        raise ValueError(f"{path}: expected {n_atoms} atoms")
    sites = []
    for r in atoms:
        eps, sig = ff.lj(r[1])
        _, q = ff.mass_charge(r[1], r[1])
        sites.append((eps, sig, q))
    name = os.path.splitext(os.path.basename(path))[0]
    if n_atoms == 1 or kind == "rigid":
        if n_groups != 1:
            raise NotImplementedError(f"{path}: molecules of more than one group")
        body = np.array([[float(v) for v in r[2:5]] if len(r) >= 5 else [0.0, 0.0, 0.0] for r in atoms])
        probe = W.Probe(name, sites, body, start=start, k_fb=k_fb, k_or=k_or)
        probe.labels = [r[1] for r in atoms]
        return probe
    if n_groups != 1 or kind != "flexible":
        raise NotImplementedError(f"{path}: only molecules of one flexible or rigid group")
    # counts: chiral centres, bonds, bond dipoles, bends, Urey-Bradley,
    # inversion bends, torsions, improper torsions, bond/bond,
    # stretch/bend, bend/bend, stretch/torsion, bend/torsion, intra VDW,
    # intra Coulomb (src/molecule.c:1445-1466)
    i = first + n_atoms
    counts = [int(v) for v in data[i][:15]]
    i += 1
    (n_chiral, n_bond, n_dip, n_bend, n_ub, n_inv, n_tors, n_imp, n_bb, n_sb, n_bb2, n_st, n_bt,
     n_intra, n_intra_cc) = counts
    if n_chiral or n_dip or n_ub or n_inv or n_imp or n_bb or n_sb or n_bb2 or n_st or n_bt:
        raise NotImplementedError(f"{path}: chiral centres, bond dipoles or cross terms")

    def section(n):
        nonlocal i
        out = data[i:i + n]
        i += n
        return out

    def kind_of(table, word):
        if word not in table:
            raise NotImplementedError(f"{path}: {word}")
        return table[word]

    bonds = [(int(r[0]), int(r[1]), kind_of(_BONDS, r[2]), *[float(v) for v in r[3:]]) for r in section(n_bond)]
    bends = [(int(r[0]), int(r[1]), int(r[2]), kind_of(_BENDS, r[3]), *[float(v) for v in r[4:]])
             for r in section(n_bend)]
    torsions = [(int(r[0]), int(r[1]), int(r[2]), int(r[3]), kind_of(_TORSIONS, r[4]),
                 *[float(v) for v in r[5:9]]) for r in section(n_tors)]
    intra = [(int(r[0]), int(r[1])) for r in section(n_intra)]
    if n_intra_cc and any(abs(q) > 1e-10 for _, _, q in sites):
        raise NotImplementedError(f"{path}: charged flexible molecules")
    probe = W.ChainProbe(name, sites, bonds, bends, torsions, intra, start=start, k_fb=k_fb,
                         k_trial=k_or, k_torsion=k_torsion, moves=moves_per_open_bead)
    probe.labels = [r[1] for r in atoms]
    return probe


# --------------------------------------- the reference wrapper's calls
def _share(raspa_dir):
    d = raspa_dir or os.environ.get("RASPA_DIR")
    if not d:
        raise ValueError("pass raspa_dir or set RASPA_DIR (the directory holding share/raspa)")
    return os.path.join(d, "share", "raspa")


def widom(structure, molecule, unit_cells=(1, 1, 1), cycles=2000, forcefield="CrystalGenerator",
          temperature=298.0, cutoff=12.8, raspa_dir=None, molecule_set="TraPPE", insertions=None,
          seed=1, charge_method=None, precision=R.PREC_MIXED, device=0, ideal_gas_weight=1.0):
    """Widom insertion of `molecule` in a CIF structure given as text, as the
    reference's scripts set it up (python/raspa2.py:345-387): CutOff 12.8 A,
    298 K, first-bead trials 10; a cycle of an empty system is 20 moves, so
    cycles -> 20 x cycles insertions unless `insertions` is given.  Flexible
    molecules grow by CBMC with the internal MC scheme.  Returns the
    estimators of src/statistics.c (average Widom Rosenbluth weight, Henry
    coefficient [mol/kg/Pa] with their 2-sigma errors, <U_gh> [K]); the Henry
    coefficient divides by ideal_gas_weight (IdealGasRosenbluthWeight, 1 unless
    set, src/input.c:902, 2906; see ideal_gas_rosenbluth_weight)."""
    share = _share(raspa_dir)
    ff = ForceField.load(os.path.join(share, "forcefield", forcefield))
    fw = framework_from_cif(structure, ff, unit_cells)
    probe = read_molecule(os.path.join(share, "molecules", molecule_set, molecule + ".def"), ff)
    ff.check_pairs(probe.labels, fw.labels)
    flexible = isinstance(probe, W.ChainProbe)
    charged = (not flexible) and any(abs(q) > 1e-10 for q in fw.charge) and probe.charged
    cm = charge_method if charge_method is not None else (R.CHARGE_EWALD if charged else R.CHARGE_NONE)
    ctx = fw.context(cutoff=cutoff, charge_method=cm, temperature=temperature, precision=precision,
                     device=device, shifted=ff.shifted)
    n = insertions if insertions is not None else 20 * cycles
    if flexible:
        _, chain = W.register_chain(ctx, probe)
        blocks, _, _ = ctx.widom_chain_run(chain, seed, n)
    else:
        _, mol = W.register_probe(ctx, probe)
        blocks, _ = ctx.widom_run(mol, seed, n)
    est = R.widom_estimators(blocks, temperature, ctx.info()["density"],
                             ideal_gas_weight=ideal_gas_weight)
    # "Average Widom Rosenbluth-weight" per block and its 2-sigma spread over
    # the five blocks (src/statistics.c:4747-4800)
    b = blocks.as_array()
    wb = np.where(b[:, 0] > 0, b[:, 1] / np.maximum(b[:, 0], 1.0), 0.0)
    est["w_blocks"] = wb
    est["w"] = float(wb.mean())
    est["w_err"] = float(2.0 * np.sqrt(abs((wb * wb).mean() - wb.mean() ** 2)))
    est["insertions"] = n
    ctx.close()
    return est


def ideal_gas_rosenbluth_weight(molecule, forcefield="CrystalGenerator", temperature=298.0, cutoff=12.8,
                                raspa_dir=None, molecule_set="TraPPE", insertions=1 << 18, seed=1,
                                device=0):
    """IdealGasRosenbluthWeight of a molecule: the average Rosenbluth weight
    of its growth with no framework and no other molecules (the value a RASPA2
    input supplies for the Henry coefficient of a flexible molecule,
    src/input.c:2906, src/mc_moves.c:7013), measured by the same growth on
    the device in an empty box."""
    share = _share(raspa_dir)
    ff = ForceField.load(os.path.join(share, "forcefield", forcefield))
    probe = read_molecule(os.path.join(share, "molecules", molecule_set, molecule + ".def"), ff)
    if not isinstance(probe, W.ChainProbe):
        return 1.0
    ff.check_pairs(probe.labels, ())
    e = np.zeros(0)
    box = np.diag([4.0 * cutoff] * 3).T.reshape(-1)
    ctx = R.Context(box, R.BC_RECTANGULAR, e, e, e, e, e, charge=e, mass=e, cutoff_vdw=cutoff,
                    shifted=ff.shifted, temperature=temperature, device=device, precision=R.PREC_FP64)
    try:
        _, chain = W.register_chain(ctx, probe)
        blocks, _, _ = ctx.widom_chain_run(chain, seed, insertions)
        b = blocks.as_array()
        return float(b[:, 1].sum() / b[:, 0].sum())
    finally:
        ctx.close()


def psd_columns(hist, cycles, range_):
    """The three columns of the reference's PSD histogram file
    (src/sample.c:3552-3575): diameter 2 delta (k + 1/2), the cumulative
    fraction hist[k] / cycles, and minus its finite-difference derivative
    (one-sided at the ends, central inside)."""
    hist = np.asarray(hist, dtype=np.float64)
    bins = len(hist)
    delta = range_ / bins
    d, mids, psd = [], [], []
    for k in range(bins):
        mid = hist[k] / cycles
        if k == 0:
            deriv = (hist[k + 1] / cycles - mid) / delta
        elif k == bins - 1:
            deriv = (mid - hist[k - 1] / cycles) / delta
        else:
            deriv = (hist[k + 1] / cycles - hist[k - 1] / cycles) / (2.0 * delta)
        d.append(2.0 * delta * (k + 0.5))
        mids.append(mid)
        psd.append(-deriv)
    return np.array(d), np.array(mids), np.array(psd)


def get_pore_size_distribution(structure, unit_cells=(1, 1, 1), cycles=500, input_file_type="cif",
                               forcefield="CrystalGenerator", bins=50, raspa_dir=None, seed=1,
                               probe_distance="Sigma", psd_range=12.5, device=0):
    """The wrapper's pore-size distribution (python/raspa2.py:390-441:
    SimulationType PSD, PSDProbeDistance Sigma, PSDRange 12.5,
    PSDHistogramSize `bins`, 20 x cycles sphere centres per cycle) on the
    device.  Returns [[d1..dn], [psd1..psdn]]: pore diameter [A] and the PSD
    column of the reference's histogram file."""
    if input_file_type.lower() != "cif":
        raise NotImplementedError("only CIF structures")
    share = _share(raspa_dir)
    ff = ForceField.load(os.path.join(share, "forcefield", forcefield))
    fw = framework_from_cif(structure, ff, unit_cells)
    factor = 2.0 ** (1.0 / 6.0) if probe_distance.lower() == "minimum" else 1.0
    ctx = fw.context(device=device, precision=R.PREC_FP64, shifted=ff.shifted)
    try:
        hist = ctx.PoreSizeDistribution(cycles, 20 * cycles, seed=seed, bins=bins, range_=psd_range,
                                        probe_factor=factor)
    finally:
        ctx.close()
    d, _, psd = psd_columns(hist, cycles, psd_range)
    return [list(d), list(psd)]


def get_helium_void_fraction(structure, unit_cells=(1, 1, 1), cycles=2000, input_file_type="cif",
                             forcefield="CrystalGenerator", raspa_dir=None, insertions=None, seed=1,
                             device=0):
    """Helium void fraction: the average Widom Rosenbluth weight of helium
    (python/raspa2.py:345-387)."""
    if input_file_type != "cif":
        raise NotImplementedError("only CIF structures")
    est = widom(structure, "helium", unit_cells, cycles, forcefield, raspa_dir=raspa_dir,
                insertions=insertions, seed=seed, precision=R.PREC_FP64, device=device)
    return est["w"]


AVOGADRO = 6.02214076e23


def get_geometric_surface_area(structure, unit_cells=(1, 1, 1), cycles=500, input_file_type="cif",
                               units="m^2/g", forcefield="CrystalGenerator", raspa_dir=None, seed=1,
                               samples=None, device=0):
    """Geometric surface area with the wrapper's settings (python/raspa2.py:
    297-342): N2's first bead as the probe, SurfaceAreaProbeDistance Sigma
    (the contact distance is the mixed sigma).  A cycle of an empty system is
    20 moves of 5 points per atom, so cycles -> 100 x cycles points per atom
    unless `samples` is given.  units: "A^2", "m^2/g" or "m^2/cm^3"."""
    if input_file_type != "cif":
        raise NotImplementedError("only CIF structures")
    share = _share(raspa_dir)
    ff = ForceField.load(os.path.join(share, "forcefield", forcefield))
    fw = framework_from_cif(structure, ff, unit_cells)
    n2 = read_molecule(os.path.join(share, "molecules", "TraPPE", "N2.def"), ff)
    ff.check_pairs(n2.labels[:1], fw.labels)
    ctx = fw.context(cutoff=12.8, temperature=298.0, device=device, shifted=ff.shifted)
    t = ctx.add_pseudo_atom(*n2.sites[n2.start])
    mask = np.array([ff.interacts(lab) for lab in fw.labels], dtype=np.uint8)
    area, _ = ctx.SurfaceArea(t, samples if samples is not None else 100 * cycles, seed=seed,
                              distance_factor=1.0, interacts=mask)
    info = ctx.info()
    ctx.close()
    if units == "A^2":
        return area
    if units == "m^2/g":
        return area * 1.0e-20 * AVOGADRO / float(np.sum(fw.mass))
    if units == "m^2/cm^3":
        return area / info["volume"] * 1.0e4
    raise ValueError(f"units {units!r}")

