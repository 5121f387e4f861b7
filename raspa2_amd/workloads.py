"""Synthetic frameworks and adsorbates of the benchmark configurations.

Every workload is deterministic in its seeds.  Positions come from
rhg_synthetic_positions (random sequential addition, minimum-image minimum
separation); LJ parameters and charges from numpy's PCG64 generator.

Cells are built with RASPA2's construction of the Box matrix from
a, b, c, alpha, beta, gamma (src/framework.c:3024-3066), with libm cos/sin.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import rhg as R

KELVIN_PER_ENERGY = None  # filled lazily from the units of a context


def cell_box(a, b, c, alpha=90.0, beta=90.0, gamma=90.0):
    """RASPA2's Box (ax ay az bx by bz cx cy cz) from cell lengths and angles."""
    al = alpha * (math.pi / 180.0)
    be = beta * (math.pi / 180.0)
    ga = gamma * (math.pi / 180.0)
    tempd = (math.cos(al) - math.cos(ga) * math.cos(be)) / math.sin(ga)
    return np.array([a, 0.0, 0.0,
                     b * math.cos(ga), b * math.sin(ga), 0.0,
                     c * math.cos(be), c * tempd,
                     c * math.sqrt(1.0 - math.cos(be) ** 2 - tempd ** 2)])


def perpendicular_widths(box):
    m = np.asarray(box, dtype=np.float64).reshape(3, 3).T  # columns = cell vectors
    a, b, c = m[:, 0], m[:, 1], m[:, 2]
    v = abs(np.dot(a, np.cross(b, c)))
    return np.array([v / np.linalg.norm(np.cross(b, c)), v / np.linalg.norm(np.cross(c, a)),
                     v / np.linalg.norm(np.cross(a, b))])


def synthetic_positions(seed, box, n, d_min, max_attempts=None):
    box = np.ascontiguousarray(box, dtype=np.float64)
    xyz = np.zeros((n, 3))
    placed = ctypes.c_int32()
    if max_attempts is None:
        max_attempts = max(10**7, 20000 * max(n, 1))
    rc = R.lib.rhg_synthetic_positions(ctypes.c_uint64(seed), box.ctypes.data_as(ctypes.c_void_p),
                                       n, d_min, max_attempts,
                                       xyz.ctypes.data_as(ctypes.c_void_p), ctypes.byref(placed))
    if rc:
        raise R.RhgError(f"could place only {placed.value} of {n} atoms at d_min={d_min}")
    return xyz


@dataclass
class FrameworkSpec:
    name: str
    box: np.ndarray
    bc: int
    xyz: np.ndarray
    eps_k: np.ndarray
    sigma: np.ndarray
    charge: np.ndarray
    mass: np.ndarray
    unit_box: np.ndarray = None

    @property
    def n(self):
        return len(self.xyz)

    def context(self, cutoff=12.0, charge_method=R.CHARGE_NONE, ewald_precision=1e-6,
                temperature=300.0, precision=R.PREC_FP64, device=0, shifted=True):
        return R.Context(self.box, self.bc, self.xyz[:, 0], self.xyz[:, 1], self.xyz[:, 2],
                         self.eps_k, self.sigma, charge=self.charge, mass=self.mass,
                         cutoff_vdw=cutoff, shifted=shifted, charge_method=charge_method,
                         cutoff_charge=cutoff, ewald_precision=ewald_precision,
                         temperature=temperature, device=device, precision=precision)

    def oracle_args(self):
        return (self.box, self.bc, self.xyz[:, 0], self.xyz[:, 1], self.xyz[:, 2], self.charge,
                self.eps_k, self.sigma, self.mass)


@dataclass
class Probe:
    """A rigid adsorbate: sites (eps_k, sigma, charge) and body positions."""
    name: str
    sites: list            # [(eps_k, sigma, charge)]
    body: np.ndarray       # n x 3
    start: int = 0
    k_fb: int = 1
    k_or: int = 10

    @property
    def n_sites(self):
        return len(self.sites)

    @property
    def charged(self):
        return any(abs(q) > 1e-10 for _, _, q in self.sites)


METHANE = Probe("methane", [(148.0, 3.73, 0.0)], np.zeros((1, 3)))
# TraPPE CO2 site order O, C, O; growth starts at the carbon
CO2 = Probe("CO2", [(79.0, 3.05, -0.35), (27.0, 2.80, 0.70), (79.0, 3.05, -0.35)],
            np.array([[0.0, 0.0, 1.16], [0.0, 0.0, 0.0], [0.0, 0.0, -1.16]]), start=1)
N2LIKE = Probe("N2", [(36.0, 3.31, 0.0), (36.0, 3.31, 0.0)],
               np.array([[0.0, 0.0, 0.55], [0.0, 0.0, -0.55]]), start=0)
HELIUM = Probe("helium", [(10.9, 2.64, 0.0)], np.zeros((1, 3)), k_fb=10)


@dataclass
class ChainProbe:
    """A flexible adsorbate of one flexible group: per-atom sites (eps_k,
    sigma, charge) and bonded terms in molecule-file units (K, A, degrees):
    bonds (a, b, kind, p0, p1), bends (a, b, c, kind, p0, p1), torsions
    (a, b, c, d, kind, p0..p3), intra-molecular LJ pairs (a, b)."""
    name: str
    sites: list
    bonds: list
    bends: list
    torsions: list
    intra: list
    start: int = 0
    k_fb: int = 10
    k_trial: int = 10
    k_torsion: int = 100
    moves: int = 150

    @property
    def n_atoms(self):
        return len(self.sites)


def trappe_alkane(n, **kw):
    """n-alkane in the TraPPE united-atom model (molecules/TraPPE/C<n>.def,
    forcefield TraPPE CH3_sp3 98 K 3.75 A, CH2_sp3 46 K 3.95 A): harmonic
    bonds 96500 K/A^2 at 1.54 A, harmonic bends 62500 K/rad^2 at 114 deg,
    TRAPPE_DIHEDRAL 0, 355.03, -68.19, 791.32 K, LJ between beads four or
    more bonds apart."""
    ch3, ch2 = (98.0, 3.75, 0.0), (46.0, 3.95, 0.0)
    sites = [ch3 if i in (0, n - 1) else ch2 for i in range(n)] if n > 1 else [(148.0, 3.73, 0.0)]
    bonds = [(i, i + 1, R.BOND_HARMONIC, 96500.0, 1.54) for i in range(n - 1)]
    bends = [(i, i + 1, i + 2, R.BEND_HARMONIC, 62500.0, 114.0) for i in range(n - 2)]
    torsions = [(i, i + 1, i + 2, i + 3, R.TORSION_TRAPPE, 0.0, 355.03, -68.19, 791.32)
                for i in range(n - 3)]
    intra = [(i, j) for i in range(n) for j in range(i + 4, n)]
    return ChainProbe(f"C{n}", sites, bonds, bends, torsions, intra, **kw)


def register_chain(ctx, probe):
    """Adds the chain's pseudo-atoms to a context and returns its rhg.Chain."""
    types = [ctx.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    return types, R.Chain(types, probe.bonds, probe.bends, probe.torsions, probe.intra,
                          start=probe.start, k_fb=probe.k_fb, k_trial=probe.k_trial,
                          k_torsion=probe.k_torsion, moves_per_open_bead=probe.moves)


def register_probe(ctx, probe):
    """Adds the probe's pseudo-atoms to a context and returns its rhg_molecule."""
    types = [ctx.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    return types, R.molecule(types, probe.body, start=probe.start, k_fb=probe.k_fb,
                             k_or=probe.k_or)


def _lj_params(seed, n):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(20.0, 80.0, n), rng.uniform(2.5, 3.5, n)


def _charges(seed, n, sd=0.4):
    rng = np.random.Generator(np.random.PCG64(seed))
    q = rng.normal(0.0, sd, n)
    return q - q.mean()


def c0_framework():
    """Cubic 24.5 A cell, 576 atoms, 2.6 A minimum separation, no charges."""
    box = cell_box(24.5, 24.5, 24.5)
    xyz = synthetic_positions(1000, box, 576, 2.6)
    eps, sig = _lj_params(1001, 576)
    return FrameworkSpec("c0", box, R.BC_RECTANGULAR, xyz, eps, sig, np.zeros(576),
                         np.full(576, 12.0), unit_box=box)


def c1_framework(n_rep=2):
    """25.8 A cubic unit cell of 576 atoms replicated n_rep^3 (4608 atoms at
    2x2x2); charges N(0, 0.4 e) shifted to zero net charge."""
    ubox = cell_box(25.8, 25.8, 25.8)
    uxyz = synthetic_positions(1100, ubox, 576, 2.6)
    ueps, usig = _lj_params(1101, 576)
    uq = _charges(1102, 576)
    box = cell_box(n_rep * 25.8, n_rep * 25.8, n_rep * 25.8)
    xyz, eps, sig, q = [], [], [], []
    m = ubox.reshape(3, 3)  # rows a, b, c
    for i in range(n_rep):
        for j in range(n_rep):
            for k in range(n_rep):
                off = i * m[0] + j * m[1] + k * m[2]
                xyz.append(uxyz + off)
                eps.append(ueps)
                sig.append(usig)
                q.append(uq)
    xyz = np.concatenate(xyz)
    n = len(xyz)
    return FrameworkSpec(f"c1_{n_rep}x{n_rep}x{n_rep}", box, R.BC_RECTANGULAR, xyz,
                         np.concatenate(eps), np.concatenate(sig), np.concatenate(q),
                         np.full(n, 12.0), unit_box=ubox)


def c2_framework(n_atoms=1200, d_min=2.0):
    """Triclinic a=b=c=26 A, alpha=beta=gamma=80 deg, 1200 atoms.  The
    minimum separation is 2.0 A: 1200 atoms at 2.6 A in 16858 A^3 would need a
    packing fraction of 0.65, above random close packing."""
    box = cell_box(26.0, 26.0, 26.0, 80.0, 80.0, 80.0)
    xyz = synthetic_positions(1200, box, n_atoms, d_min)
    eps, sig = _lj_params(1201, n_atoms)
    return FrameworkSpec("c2", box, R.BC_TRICLINIC, xyz, eps, sig, np.zeros(n_atoms),
                         np.full(n_atoms, 12.0), unit_box=box)


@dataclass
class ScreeningJob:
    index: int
    framework: FrameworkSpec
    probe: Probe
    n_trials: int
    seed: int


def c3_job(i, min_width=24.0):
    """Screening framework i: orthorhombic edges U[12,40] A replicated so that
    every perpendicular width is >= 24 A, density U[0.03,0.06] atoms/A^3,
    N = round(density*V) clamped to [100, 8000], atoms placed directly in the
    supercell with a minimum separation min(2.6, 0.8*density^-1/3) A, charges
    as in the CO2 configuration; methane or CO2 with equal probability;
    10^4..10^6 trials, log-uniform."""
    rng = np.random.Generator(np.random.PCG64(1300 + i))
    edges = rng.uniform(12.0, 40.0, 3)
    reps = np.ceil(min_width / edges).astype(int)
    L = edges * reps
    rho = rng.uniform(0.03, 0.06)
    n = int(np.clip(round(rho * L.prod()), 100, 8000))
    rho_eff = n / L.prod()
    d_min = min(2.6, 0.8 * rho_eff ** (-1.0 / 3.0))
    box = cell_box(L[0], L[1], L[2])
    xyz = synthetic_positions(10_000 + i, box, n, d_min)
    eps = rng.uniform(20.0, 80.0, n)
    sig = rng.uniform(2.5, 3.5, n)
    q = rng.normal(0.0, 0.4, n)
    q -= q.mean()
    probe = METHANE if rng.uniform() < 0.5 else CO2
    n_trials = int(round(10 ** rng.uniform(4.0, 6.0)))
    fw = FrameworkSpec(f"c3_{i}", box, R.BC_RECTANGULAR, xyz, eps, sig, q, np.full(n, 12.0),
                       unit_box=box)
    return ScreeningJob(i, fw, probe, n_trials, 20_000 + i)


def block_pocket_centres(fractional, radii, box, n_cells=(1, 1, 1), shift=(0.0, 0.0, 0.0)):
    """Blocked-pocket centres of a system of n_cells unit cells, as
    src/framework.c:6209-6226 builds them: for each pocket (fractional
    coordinates of the unit cell plus the unit-cell shift) and each unit cell
    (j, k, l), the point box @ ((f + (j, k, l)) / n_cells).  Returns the
    centres (xyz) and their radii in that order."""
    f = np.asarray(fractional, dtype=np.float64).reshape(-1, 3)
    B = np.asarray(box, dtype=np.float64).reshape(3, 3)  # rows a, b, c
    cen, rad = [], []
    for i in range(len(f)):
        t = f[i] + np.asarray(shift)
        for j in range(n_cells[0]):
            for k in range(n_cells[1]):
                for l in range(n_cells[2]):
                    v = np.array([(t[0] + j) / n_cells[0], (t[1] + k) / n_cells[1],
                                  (t[2] + l) / n_cells[2]])
                    cen.append(v @ B)
                    rad.append(radii[i])
    return np.array(cen), np.array(rad)


def irmof1_dubbeldam(golden_dir):
    """IRMOF-1 (the P1 expansion of the reference's CIF, one unit cell) with
    the Dubbeldam2007FlexibleIRMOF-1 parameters and charges, the TraPPE CO2
    probe (growth from bead 0, RASPA2's default) and the settings of the
    reference's MakingGrids/UsingGrids example decks, read from the fixtures
    irmof1.json and irmof1_dubbeldam.json in golden_dir.  Returns
    (FrameworkSpec, Probe, deck settings)."""
    import json
    import os
    G = json.load(open(os.path.join(golden_dir, "irmof1.json")))
    D = json.load(open(os.path.join(golden_dir, "irmof1_dubbeldam.json")))
    box = cell_box(*G["cell"])
    frac = np.array([a["frac"] for a in G["atoms"]])
    xyz = frac @ box.reshape(3, 3)
    T = D["framework_types"]
    lab = [a["label"] for a in G["atoms"]]
    fw = FrameworkSpec("IRMOF-1", box, R.BC_RECTANGULAR, xyz,
                       np.array([T[l]["eps_k"] for l in lab]), np.array([T[l]["sigma"] for l in lab]),
                       np.array([T[l]["charge"] for l in lab]), np.array([T[l]["mass"] for l in lab]),
                       unit_box=box)
    P = D["probe_types"]
    probe = Probe("CO2", [(P[n]["eps_k"], P[n]["sigma"], P[n]["charge"]) for n, _ in D["co2"]],
                  np.array([p for _, p in D["co2"]]), start=0)
    return fw, probe, D
