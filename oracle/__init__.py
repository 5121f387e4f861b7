"""ctypes binding of the oracle (CPU FP64 restatement of RASPA2's host-guest path).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, where it is the checker.  The product package
(raspa2_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

BC_RECTANGULAR = 1
BC_TRICLINIC = 2
CHARGE_NONE = 0
CHARGE_EWALD = 1
MAX_SITES = 8

REC = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
           FBMASK=9, ORMASK=10, W1=11, W2=12, STAGE=13)
REC_N = 16


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        _declare(_lib)
    return _lib


_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_P = ctypes.c_void_p


class Molecule(ctypes.Structure):
    _fields_ = [("n_sites", _I), ("type", _I * MAX_SITES), ("body", (_D * 3) * MAX_SITES),
                ("start", _I), ("k_fb", _I), ("k_or", _I)]


class Trials(ctypes.Structure):
    _fields_ = [("fb", (_D * 3) * 64), ("cord", ((_D * 3) * MAX_SITES) * 64),
                ("u_sel_fb", _D), ("u_sel_or", _D)]


CHAIN_MAX = 64
CHAIN_MAX_INTRA = 2016
CREC = dict(W=0, DU=1, UHOSTVDW=2, UINTRA=3, UBOND=4, UBEND=5, UTORSION=6, OVERLAP=7, STAGE=8,
            STEP=9, W1=10)
CREC_N = 11
BOND_FIXED = 0
BOND_HARMONIC = 1
BEND_HARMONIC = 0
TORSION_TRAPPE = 0


class Chain(ctypes.Structure):
    _fields_ = [("n_atoms", _I), ("start", _I), ("type", _I * CHAIN_MAX),
                ("n_bonds", _I), ("bond", (_I * 3) * CHAIN_MAX), ("bond_par", (_D * 4) * CHAIN_MAX),
                ("n_bends", _I), ("bend", (_I * 4) * (2 * CHAIN_MAX)),
                ("bend_par", (_D * 4) * (2 * CHAIN_MAX)),
                ("n_torsions", _I), ("torsion", (_I * 5) * (4 * CHAIN_MAX)),
                ("torsion_par", (_D * 4) * (4 * CHAIN_MAX)),
                ("n_intra", _I), ("intra", (_I * 2) * CHAIN_MAX_INTRA), ("intra14_scaling", _D),
                ("k_fb", _I), ("k_trial", _I), ("k_torsion", _I), ("moves", _I),
                ("max_bond", _D * CHAIN_MAX), ("max_bend", _D * CHAIN_MAX), ("max_cone", _D * CHAIN_MAX)]


def _declare(L):
    def f(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)
    f("orc_units", None, _P)
    f("orc_mt64_sequence", None, _U64, _I64, _P)
    f("orc_mt64_uniform_sequence", None, _U64, _I64, _P)
    f("orc_stream_key", _U64, _U64, _U64)
    f("orc_stream_uniform", _D, _U64, _U64)
    f("orc_stream_uniforms", None, _U64, _U64, _I64, _P)
    f("orc_cell_make", None, _D, _D, _D, _D, _D, _D, _P, _P, _P)
    f("orc_invert3x3", None, _P, _P, _P)
    f("orc_cell_properties", None, _P, _P)
    f("orc_min_image", None, _P, _P, _I, _P)
    f("orc_system_create", _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D, _I, _I, _D, _D, _D)
    f("orc_system_destroy", None, _P)
    f("orc_add_pseudo_atom", _I, _P, _D, _D, _D)
    f("orc_pair_params", None, _P, _I, _I, _P)
    f("orc_system_info", None, _P, _P)
    f("orc_vdw_energy_at", _D, _P, _I, _P, _D)
    f("orc_charge_energy_at", _D, _P, _I, _P, _D)
    f("orc_energy_at_many", None, _P, _I, _I64, _P, _P, _P)
    f("orc_ewald_nk", _I, _P)
    f("orc_ewald_tables", None, _P, _P, _P, _P, _P, _P)
    f("orc_ewald_fourier_self_energy_framework", _D, _P)
    f("orc_ewald_real_energy_framework", _D, _P)
    f("orc_ewald_self_energy_framework", _D, _P)
    f("orc_fourier_adsorbate", None, _P, _I, _P, _P, _P)
    f("orc_widom_insertion_trials", None, _P, _P, _P, _P)
    f("orc_widom_retrace_trials", None, _P, _P, _I, _P, _P, _P)
    f("orc_widom_run_stream", None, _P, _P, _U64, _I64, _I64, _P)
    f("orc_stream_trials", None, _P, _P, _U64, _I64, _P, _P, _P)
    f("orc_widom_run_mt", _I64, _P, _P, _U64, _I64, _I, _P)
    f("orc_widom_run_mt_trials", _I64, _P, _P, _U64, _I64, _I, _P, _I64, _P, _P, _P, _P)
    f("orc_block_of_cycles", None, _I64, _P)
    f("orc_widom_estimators", None, _P, _D, _D, _D, _P)
    f("orc_framework_density", _D, _D, _D)
    f("orc_grid_geometry", None, _P, _D, _P)
    f("orc_derivatives_vdw_at", None, _P, _I, _P, _P)
    f("orc_grid_node", None, _P, _I, _P, _I, _I, _I, _P)
    f("orc_lm_coeff", None, _P)
    f("orc_interpolate_grid", _D, _P, _P, _I, _P, _P, _P)
    f("orc_interpolate_coulomb_grid", _D, _P, _P, _I, _P, _P, _P, _D)
    f("orc_interpolate_direct", _D, _P, _I, _P, _I, _P, _P, _P)
    f("orc_set_block_pockets", _I, _P, _I, _P, _P, _P, _I)
    f("orc_derivatives_real_at", _D, _P, _P, _P)
    f("orc_use_grid", _I, _P, _I, _P, _P, _P)
    f("orc_set_adsorbates", _I, _P, _I, _P, _P, _P)
    f("orc_inter_vdw_at", _D, _P, _I, _P, _I, _D)
    f("orc_inter_charge_at", _D, _P, _I, _P, _I, _D)
    f("orc_fourier_move", None, _P, _I, _P, _P, _I, _P)
    f("orc_surface_area", _D, _P, _I, _D, _I64, _U64, _P, _P)
    f("orc_use_coulomb_grid", _I, _P, _P, _P, _P)
    f("orc_coulomb_grid_node", None, _P, _P, _I, _I, _I, _P)
    f("orc_blocked_pocket", _I, _P, _P)
    f("orc_psd_largest", _D, _P, _D, _I64, _U64, _I64)
    f("orc_psd_histogram", None, _P, _D, _I64, _U64, _I64, _I64, _I, _D, _P)
    f("orc_chain_plan", _I, _P, _P, _P)
    f("orc_chain_run_stream", None, _P, _P, _U64, _I64, _I64, _P, _P)
    f("orc_chain_run_mt", None, _P, _P, _U64, _I64, _P, _P)
    f("orc_chain_substream", _U64, _U64, _U64)
    f("orc_chain_bond_energy", _D, _P, _I, _P, _D)
    f("orc_chain_bend_energy", _D, _P, _I, _P, _P, _P, _P)
    f("orc_chain_torsion_energy", _D, _P, _I, _P, _P, _P, _P, _P)


def _ptr(a):
    return a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def units():
    out = np.zeros(6)
    lib().orc_units(_ptr(out))
    return dict(energy=out[0], coulomb=out[1], energy_to_kelvin=out[2], kelvin_to_energy=out[3],
                k_b=out[4], coulomb_kelvin=out[5])


def mt64(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_mt64_sequence(seed, n, _ptr(out))
    return out


def mt64_uniform(seed, n):
    out = np.zeros(n)
    lib().orc_mt64_uniform_sequence(seed, n, _ptr(out))
    return out


def stream_uniforms(seed, insertion, n):
    out = np.zeros(n)
    lib().orc_stream_uniforms(seed, insertion, n, _ptr(out))
    return out


def cell_make(a, b, c, alpha=90.0, beta=90.0, gamma=90.0):
    box, inv, props = np.zeros(9), np.zeros(9), np.zeros(7)
    lib().orc_cell_make(a, b, c, alpha, beta, gamma, _ptr(box), _ptr(inv), _ptr(props))
    return box, inv, props


def invert3x3(m):
    m = _f64(m)
    out = np.zeros(9)
    det = np.zeros(1)
    lib().orc_invert3x3(_ptr(m), _ptr(out), _ptr(det))
    return out, det[0]


def min_image(box, inv, bc, dr):
    box, inv = _f64(box), _f64(inv)
    d = _f64(dr).copy()
    lib().orc_min_image(_ptr(box), _ptr(inv), bc, _ptr(d))
    return d


class System:
    """One framework + force field + (optional) Ewald setup."""

    def __init__(self, box, bc, x, y, z, charge, eps_k, sigma, mass=None, cutoff_vdw=12.0,
                 shifted=True, charge_method=CHARGE_NONE, cutoff_charge=12.0,
                 ewald_precision=1e-6, temperature=300.0):
        L = lib()
        self._keep = [_f64(box), _f64(x), _f64(y), _f64(z), _f64(charge), _f64(eps_k), _f64(sigma)]
        self._keep.append(_f64(mass) if mass is not None else None)
        b, xx, yy, zz, qq, ee, ss, mm = self._keep
        self.n = len(xx)
        self.temperature = temperature
        self.h = L.orc_system_create(_ptr(b), bc, self.n, _ptr(xx), _ptr(yy), _ptr(zz), _ptr(qq),
                                     _ptr(ee), _ptr(ss), _ptr(mm) if mm is not None else None,
                                     cutoff_vdw, 1 if shifted else 0, charge_method,
                                     cutoff_charge, ewald_precision, temperature)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_system_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def add_pseudo_atom(self, eps_k, sigma, charge=0.0):
        t = lib().orc_add_pseudo_atom(self.h, eps_k, sigma, charge)
        if not hasattr(self, "_tq"):
            self._tq = {}
            self._pseudo = {}
        self._tq[t] = float(charge)
        self._pseudo[t] = (float(eps_k), float(sigma))
        return t

    def t_q(self, t):
        return self._tq[t]

    def pair_params(self, t, atom):
        out = np.zeros(3)
        lib().orc_pair_params(self.h, t, atom, _ptr(out))
        return out

    def info(self):
        out = np.zeros(8)
        lib().orc_system_info(self.h, _ptr(out))
        return dict(alpha=out[0], kx=int(out[1]), ky=int(out[2]), kz=int(out[3]), nk=int(out[4]),
                    uion=out[5], beta=out[6], volume=out[7])

    def vdw_at(self, t, pos, scaling=1.0):
        p = _f64(pos)
        return lib().orc_vdw_energy_at(self.h, t, _ptr(p), scaling)

    def charge_at(self, t, pos, scaling=1.0):
        p = _f64(pos)
        return lib().orc_charge_energy_at(self.h, t, _ptr(p), scaling)

    def energy_at_many(self, t, xyz):
        xyz = _f64(xyz).reshape(-1, 3)
        n = len(xyz)
        uv, uc = np.zeros(n), np.zeros(n)
        lib().orc_energy_at_many(self.h, t, n, _ptr(xyz), _ptr(uv), _ptr(uc))
        return uv, uc

    def ewald_tables(self):
        nk = lib().orc_ewald_nk(self.h)
        kv, kf, sr, si = np.zeros(3 * nk), np.zeros(nk), np.zeros(nk), np.zeros(nk)
        ki = np.zeros(3 * nk, dtype=np.int32)
        lib().orc_ewald_tables(self.h, _ptr(kv), _ptr(kf), _ptr(sr), _ptr(si), _ptr(ki))
        return dict(kvec=kv.reshape(-1, 3), kfactor=kf, sf_re=sr, sf_im=si, kidx=ki.reshape(-1, 3))

    def fourier_self_framework(self):
        return lib().orc_ewald_fourier_self_energy_framework(self.h)

    def real_energy_framework(self):
        return lib().orc_ewald_real_energy_framework(self.h)

    def self_energy_framework(self):
        return lib().orc_ewald_self_energy_framework(self.h)

    def fourier_adsorbate(self, types, xyz):
        types = np.ascontiguousarray(types, dtype=np.int32)
        xyz = _f64(xyz).reshape(-1)
        out = np.zeros(5)
        lib().orc_fourier_adsorbate(self.h, len(types), _ptr(types), _ptr(xyz), _ptr(out))
        return out

    def derivatives_vdw_at(self, t, pos):
        p = _f64(pos)
        out = np.zeros(14)
        lib().orc_derivatives_vdw_at(self.h, t, _ptr(p), _ptr(out))
        return out

    def set_block_pockets(self, centres, radii, unit_box, invert=False):
        c = _f64(centres).reshape(-1, 3)
        r = _f64(radii).reshape(-1)
        self._pockets = (c, r)
        lib().orc_set_block_pockets(self.h, len(r), _ptr(c), _ptr(r), _ptr(_f64(unit_box)),
                                    1 if invert else 0)

    def BlockedPocket(self, pos):
        return bool(lib().orc_blocked_pocket(self.h, _ptr(_f64(pos))))

    def interpolate_direct(self, t, geom, bc, unit_box, unit_inv, pos):
        """InterpolateVDWGrid at pos with the corner nodes computed on the fly."""
        geom, ub, ui, p = _f64(geom), _f64(unit_box), _f64(unit_inv), _f64(pos)
        return lib().orc_interpolate_direct(self.h, t, _ptr(geom), bc, _ptr(ub), _ptr(ui), _ptr(p))

    def derivatives_real_at(self, pos):
        """(out[14], smallest_r) of the real-space potential of a unit charge."""
        out = np.zeros(14)
        r = lib().orc_derivatives_real_at(self.h, _ptr(_f64(pos)), _ptr(out))
        return out, r

    def set_adsorbates(self, start, types, xyz):
        """Adsorbate molecules present: molecule m holds atoms
        [start[m], start[m+1]) of the given pseudo-atom types at xyz."""
        st = np.ascontiguousarray(start, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        xy = _f64(xyz).reshape(-1, 3)
        self._ads = (st, ty, xy)
        n_mol = len(st) - 1
        assert lib().orc_set_adsorbates(self.h, n_mol, _ptr(st), _ptr(ty), _ptr(xy)) == 0

    def inter_vdw_at(self, t, pos, exclude=-1):
        return lib().orc_inter_vdw_at(self.h, t, _ptr(_f64(pos)), exclude, 1.0)

    def inter_charge_at(self, t, pos, exclude=-1, charge=None):
        q = self.t_q(t) if charge is None else charge
        return lib().orc_inter_charge_at(self.h, t, _ptr(_f64(pos)), exclude, q)

    def fourier_move(self, new_types=(), new_xyz=None, old_mol=-1):
        """CalculateEwaldFourierAdsorbate(New, Old, mol): (host, ads, self,
        excluded, sum) for a new molecule and/or removing adsorbate old_mol."""
        out = np.zeros(5)
        ty = np.ascontiguousarray(list(new_types), dtype=np.int32)
        xy = _f64(new_xyz if new_xyz is not None else np.zeros((0, 3))).reshape(-1, 3)
        lib().orc_fourier_move(self.h, len(ty), _ptr(ty) if len(ty) else None, _ptr(xy) if len(ty) else None,
                               old_mol, _ptr(out))
        return out

    def surface_area(self, t, samples, seed=1, factor=2.0 ** (1.0 / 6.0), interacts=None):
        counts = np.zeros(len(self._keep[1]), dtype=np.uint64)
        mask = None if interacts is None else np.ascontiguousarray(interacts, dtype=np.uint8)
        area = lib().orc_surface_area(self.h, t, factor, samples, seed,
                                      None if mask is None else _ptr(mask), _ptr(counts))
        return area, counts

    def use_grid(self, t, grid, geom, unit_box):
        """Energies of type t at a position come from the grid (None
        detaches); the arrays are kept alive by this object."""
        if not hasattr(self, "_grids"):
            self._grids = {}
        if grid is None:
            self._grids.pop(t, None)
            lib().orc_use_grid(self.h, t, None, None, None)
            return
        g = (np.ascontiguousarray(grid, dtype=np.float32), _f64(geom), _f64(unit_box))
        self._grids[t] = g
        assert lib().orc_use_grid(self.h, t, _ptr(g[0]), _ptr(g[1]), _ptr(g[2])) == 0

    def use_coulomb_grid(self, grid, geom, unit_box):
        if not hasattr(self, "_grids"):
            self._grids = {}
        if grid is None:
            self._grids.pop("coulomb", None)
            lib().orc_use_coulomb_grid(self.h, None, None, None)
            return
        g = (np.ascontiguousarray(grid, dtype=np.float32), _f64(geom), _f64(unit_box))
        self._grids["coulomb"] = g
        lib().orc_use_coulomb_grid(self.h, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]))

    def coulomb_grid_node(self, geom, i, j, k):
        out = np.zeros(8, dtype=np.float32)
        lib().orc_coulomb_grid_node(self.h, _ptr(_f64(geom)), i, j, k, _ptr(out))
        return out

    def grid_node(self, t, geom, i, j, k):
        g = _f64(geom)
        out = np.zeros(8, dtype=np.float32)
        lib().orc_grid_node(self.h, t, _ptr(g), i, j, k, _ptr(out))
        return out


def molecule(types, body, start=0, k_fb=10, k_or=10):
    m = Molecule()
    m.n_sites = len(types)
    for i, t in enumerate(types):
        m.type[i] = int(t)
        for d in range(3):
            m.body[i][d] = float(body[i][d])
    m.start = start
    m.k_fb = k_fb
    m.k_or = k_or
    return m


def widom_run_stream(system, mol, seed, first, n):
    rec = np.zeros((n, REC_N))
    lib().orc_widom_run_stream(system.h, ctypes.byref(mol), seed, first, n, _ptr(rec))
    return rec


def stream_trials(system, mol, seed, insertion):
    fb = np.zeros((mol.k_fb, 3))
    cord = np.zeros((mol.k_or, mol.n_sites, 3))
    usel = np.zeros(2)
    lib().orc_stream_trials(system.h, ctypes.byref(mol), seed, insertion, _ptr(fb), _ptr(cord),
                            _ptr(usel))
    return fb, cord, usel


def widom_retrace_trials(system, mol, exclude, old, fb, cord):
    """RetraceMolecule(CBMC_DELETION) of the molecule at `old` (n_sites x 3)
    with the other trials fb, cord (entries 0 ignored)."""
    t = Trials()
    for i in range(len(fb)):
        for d in range(3):
            t.fb[i][d] = float(fb[i][d])
    for iu in range(cord.shape[0]):
        for j in range(cord.shape[1]):
            for d in range(3):
                t.cord[iu][j][d] = float(cord[iu][j][d])
    rec = np.zeros(REC_N)
    o = _f64(old).reshape(-1)
    lib().orc_widom_retrace_trials(system.h, ctypes.byref(mol), exclude, _ptr(o), ctypes.byref(t), _ptr(rec))
    return rec


def widom_insertion_trials(system, mol, fb, cord, usel):
    t = Trials()
    for i in range(len(fb)):
        for d in range(3):
            t.fb[i][d] = float(fb[i][d])
    for iu in range(cord.shape[0]):
        for j in range(cord.shape[1]):
            for d in range(3):
                t.cord[iu][j][d] = float(cord[iu][j][d])
    t.u_sel_fb = float(usel[0])
    t.u_sel_or = float(usel[1])
    rec = np.zeros(REC_N)
    lib().orc_widom_insertion_trials(system.h, ctypes.byref(mol), ctypes.byref(t), _ptr(rec))
    return rec


def widom_run_mt(system, mol, seed, n_cycles, steps_per_cycle=20):
    blocks = np.zeros((5, 4))
    n = lib().orc_widom_run_mt(system.h, ctypes.byref(mol), seed, n_cycles, steps_per_cycle,
                               _ptr(blocks))
    return n, blocks


def widom_run_mt_trials(system, mol, seed, n_cycles, steps_per_cycle=20, max_ins=None):
    """RASPA2's MT19937-64 driver (orc_widom_run_mt) that also returns, for its
    first max_ins insertions, the records and the trials each drew, laid out as
    host-supplied trials: (n, blocks, rec, fb, cord, u_sel)."""
    if max_ins is None:
        max_ins = n_cycles * steps_per_cycle
    nsites = mol.n_sites
    blocks = np.zeros((5, 4))
    rec = np.zeros((max_ins, REC_N))
    fb = np.zeros((max_ins, mol.k_fb, 3))
    cord = np.zeros((max_ins, mol.k_or, nsites, 3)) if nsites > 1 else np.zeros((max_ins, 1, 1, 3))
    usel = np.zeros((max_ins, 2))
    n = lib().orc_widom_run_mt_trials(system.h, ctypes.byref(mol), seed, n_cycles, steps_per_cycle,
                                      _ptr(blocks), max_ins, _ptr(rec), _ptr(fb),
                                      _ptr(cord) if nsites > 1 else None, _ptr(usel))
    k = min(n, max_ins)
    return n, blocks, rec[:k], fb[:k], cord[:k], usel[:k]


def block_of_cycles(n_cycles):
    out = np.zeros(n_cycles, dtype=np.int32)
    lib().orc_block_of_cycles(n_cycles, _ptr(out))
    return out


def widom_estimators(blocks, temperature, density, ideal_gas_weight=1.0):
    b = _f64(blocks).reshape(-1)
    out = np.zeros(16)
    lib().orc_widom_estimators(_ptr(b), temperature, density, ideal_gas_weight, _ptr(out))
    return dict(kh_blocks=out[0:5], kh=out[5], kh_err=out[6], u_blocks=out[7:12], u=out[12],
                u_err=out[13], kh_overall=out[14], w_overall=out[15])


def framework_density(mass_g_mol, volume):
    return lib().orc_framework_density(mass_g_mol, volume)


def grid_geometry(unit_box, spacing):
    ub = _f64(unit_box)
    out = np.zeros(12)
    lib().orc_grid_geometry(_ptr(ub), spacing, _ptr(out))
    return out


def lm_coeff():
    out = np.zeros(64 * 64, dtype=np.int32)
    lib().orc_lm_coeff(_ptr(out))
    return out.reshape(64, 64)


def interpolate_grid(grid, geom, bc, unit_box, unit_inv, pos):
    g = np.ascontiguousarray(grid, dtype=np.float32)
    geom, ub, ui, p = _f64(geom), _f64(unit_box), _f64(unit_inv), _f64(pos)
    return lib().orc_interpolate_grid(_ptr(g), _ptr(geom), bc, _ptr(ub), _ptr(ui), _ptr(p))


def interpolate_coulomb_grid(grid, geom, bc, unit_box, unit_inv, pos, charge):
    g = np.ascontiguousarray(grid, dtype=np.float32)
    geom, ub, ui, p = _f64(geom), _f64(unit_box), _f64(unit_inv), _f64(pos)
    return lib().orc_interpolate_coulomb_grid(_ptr(g), _ptr(geom), bc, _ptr(ub), _ptr(ui), _ptr(p),
                                              float(charge))


# ------------------------------------------------------------ flexible chains
def chain(types, bonds=(), bends=(), torsions=(), intra=(), start=0, k_fb=10, k_trial=10,
          k_torsion=100, moves=150, intra14_scaling=1.0, max_bond=0.3, max_bend=0.3, max_cone=0.3):
    """A flexible molecule of one group.  bonds: (a, b, kind, p0, p1...),
    bends: (a, b, c, kind, p0, p1), torsions: (a, b, c, d, kind, p0..p3),
    intra: (a, b) pairs; parameters in molecule-file units (K, A, degrees)."""
    c = Chain()
    c.n_atoms = len(types)
    for i, t in enumerate(types):
        c.type[i] = int(t)
    c.start = start
    c.n_bonds = len(bonds)
    for j, b in enumerate(bonds):
        c.bond[j][0], c.bond[j][1], c.bond[j][2] = int(b[0]), int(b[1]), int(b[2])
        for q, v in enumerate(b[3:]):
            c.bond_par[j][q] = float(v)
    c.n_bends = len(bends)
    for j, b in enumerate(bends):
        for q in range(4):
            c.bend[j][q] = int(b[q])
        for q, v in enumerate(b[4:]):
            c.bend_par[j][q] = float(v)
    c.n_torsions = len(torsions)
    for j, t in enumerate(torsions):
        for q in range(5):
            c.torsion[j][q] = int(t[q])
        for q, v in enumerate(t[5:]):
            c.torsion_par[j][q] = float(v)
    c.n_intra = len(intra)
    for j, p in enumerate(intra):
        c.intra[j][0], c.intra[j][1] = int(p[0]), int(p[1])
    c.intra14_scaling = intra14_scaling
    c.k_fb, c.k_trial, c.k_torsion, c.moves = k_fb, k_trial, k_torsion, moves
    for i in range(len(types)):
        c.max_bond[i] = max_bond
        c.max_bend[i] = max_bend
        c.max_cone[i] = max_cone
    return c


def chain_plan(c):
    info = np.zeros((CHAIN_MAX, 8), dtype=np.int32)
    place = np.zeros((CHAIN_MAX, CHAIN_MAX), dtype=np.int32)
    ns = lib().orc_chain_plan(ctypes.byref(c), _ptr(info), _ptr(place))
    if ns < 0:
        return None
    return [dict(cur=int(info[i, 0]), prev=int(info[i, 1]), n_prev=int(info[i, 2]),
                 place=[int(v) for v in place[i, :info[i, 3]]], n_bonds=int(info[i, 4]),
                 n_bends=int(info[i, 5]), n_torsions=int(info[i, 6]), n_intra=int(info[i, 7]))
            for i in range(ns)]


def chain_run_stream(system, c, seed, first, n, positions=False):
    rec = np.zeros((n, CREC_N))
    xyz = np.zeros((n, c.n_atoms, 3)) if positions else None
    lib().orc_chain_run_stream(system.h, ctypes.byref(c), seed, first, n, _ptr(rec),
                               _ptr(xyz) if positions else None)
    return (rec, xyz) if positions else rec


def chain_run_mt(system, c, seed, n, positions=False):
    rec = np.zeros((n, CREC_N))
    xyz = np.zeros((n, c.n_atoms, 3)) if positions else None
    lib().orc_chain_run_mt(system.h, ctypes.byref(c), seed, n, _ptr(rec), _ptr(xyz) if positions else None)
    return (rec, xyz) if positions else rec


def chain_substream(key, tag):
    return lib().orc_chain_substream(key, tag)


def chain_bond_energy(system, kind, par, r):
    p = _f64(list(par) + [0.0] * (4 - len(par)))
    return lib().orc_chain_bond_energy(system.h, kind, _ptr(p), r)


def chain_bend_energy(system, par, a, b, c_):
    p = _f64(list(par) + [0.0] * (4 - len(par)))
    aa, bb, cc = _f64(a), _f64(b), _f64(c_)
    return lib().orc_chain_bend_energy(system.h, 0, _ptr(p), _ptr(aa), _ptr(bb), _ptr(cc))


def chain_torsion_energy(system, par, a, b, c_, d):
    p = _f64(par)
    aa, bb, cc, dd = _f64(a), _f64(b), _f64(c_), _f64(d)
    return lib().orc_chain_torsion_energy(system.h, 0, _ptr(p), _ptr(aa), _ptr(bb), _ptr(cc), _ptr(dd))


# ------------------------------------------------------ pore-size distribution
def psd_largest(system, factor, n_b, seed, i):
    return lib().orc_psd_largest(system.h, factor, n_b, seed, i)


def psd_histogram(system, factor, n_b, seed, first, n, bins, range_):
    hist = np.zeros(bins)
    lib().orc_psd_histogram(system.h, factor, n_b, seed, first, n, bins, range_, _ptr(hist))
    return hist
