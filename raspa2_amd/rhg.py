"""ctypes binding of librhg.so (include/rhg.h).

The names mirror the RASPA2 routines each entry point replaces
(src/framework_energy.h:43-46, src/ewald.h:86-94, src/grids.h:65-70,
src/framework_hessian.h:43).  Every call raises RhgError with the library's
message when the C ABI returns a non-zero status.  There is no CPU fallback:
if librhg.so is missing the import of this module fails.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RHG_LIBRARY", os.path.join(_HERE, "librhg.so"))

BC_RECTANGULAR = 1
BC_TRICLINIC = 2
CHARGE_NONE = 0
CHARGE_EWALD = 1
PREC_FP64 = 0
PREC_MIXED = 1
MAX_SITES = 8
ERR_INVALID, ERR_HIP, ERR_OOM, ERR_NOT_ADMITTED, ERR_UNSUPPORTED, ERR_FAULT_INJECTED, ERR_IO = (
    -1, -2, -3, -4, -5, -6, -7)
NR_BLOCKS = 5
REC_N = 16
REC = dict(W=0, DU=1, UVDW=2, UREAL=3, UFHOST=4, UFADS=5, OVERLAP=6, SELFB=7, SELOR=8,
           FBMASK=9, ORMASK=10, W1=11, W2=12, STAGE=13, UADSVDW=14, UADSREAL=15)

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_P = ctypes.c_void_p
_PD = ctypes.POINTER(ctypes.c_double)


class RhgError(RuntimeError):
    """A failed librhg call; `code` is the RHG_ERR_* status."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class Framework(ctypes.Structure):
    _fields_ = [("n_atoms", _I32), ("x", _PD), ("y", _PD), ("z", _PD), ("charge", _PD),
                ("eps_k", _PD), ("sigma", _PD), ("mass", _PD)]


class CellT(ctypes.Structure):
    _fields_ = [("box", _D * 9), ("boundary", _I32)]


class ForceField(ctypes.Structure):
    _fields_ = [("cutoff_vdw", _D), ("shift_vdw", _I32), ("charge_method", _I32),
                ("cutoff_charge", _D), ("ewald_precision", _D), ("temperature", _D)]


class MoleculeT(ctypes.Structure):
    _fields_ = [("n_sites", _I32), ("type", _I32 * MAX_SITES), ("body", (_D * 3) * MAX_SITES),
                ("start", _I32), ("k_fb", _I32), ("k_or", _I32)]


class ChainT(ctypes.Structure):
    """rhg_chain: a flexible molecule of one flexible group."""
    _fields_ = [("n_atoms", _I32), ("start", _I32), ("type", _P),
                ("n_bonds", _I32), ("bonds", _P), ("bond_par", _P),
                ("n_bends", _I32), ("bends", _P), ("bend_par", _P),
                ("n_torsions", _I32), ("torsions", _P), ("torsion_par", _P),
                ("n_intra_vdw", _I32), ("intra_vdw", _P), ("intra14_vdw_scaling", _D),
                ("k_fb", _I32), ("k_trial", _I32), ("k_torsion", _I32), ("moves_per_open_bead", _I32),
                ("max_bond_change", _P), ("max_bend_change", _P), ("max_cone_rotation", _P)]


BOND_FIXED = 0
BOND_HARMONIC = 1
BEND_HARMONIC = 0
TORSION_TRAPPE = 0
CREC = dict(W=0, DU=1, UHOSTVDW=2, UINTRA=3, UBOND=4, UBEND=5, UTORSION=6, OVERLAP=7, STAGE=8,
            STEP=9, W1=10)
CREC_N = 11


class Chain:
    """A flexible molecule (ReadComponentDefinition, src/molecule.c:1430-3030)
    with its CBMC settings.  bonds: (a, b, kind, p0, p1), bends: (a, b, c,
    kind, p0, p1), torsions: (a, b, c, d, kind, p0, p1, p2, p3), intra: (a, b)
    pairs; parameters in molecule-file units (K, A, degrees)."""

    def __init__(self, types, bonds=(), bends=(), torsions=(), intra=(), start=0, k_fb=10,
                 k_trial=10, k_torsion=100, moves_per_open_bead=150, intra14_scaling=1.0,
                 max_bond_change=0.3, max_bend_change=0.3, max_cone_rotation=0.3):
        n = len(types)
        self.n_atoms = n
        self.types = np.ascontiguousarray(types, dtype=np.int32)

        def ints(rows, width):
            a = np.zeros((max(len(rows), 1), width), dtype=np.int32)
            for j, r in enumerate(rows):
                a[j] = [int(v) for v in r[:width]]
            return a

        def pars(rows, skip):
            a = np.zeros((max(len(rows), 1), 4))
            for j, r in enumerate(rows):
                v = [float(x) for x in r[skip:]]
                a[j, :len(v)] = v
            return a

        self.bonds, self.bond_par = ints(bonds, 3), pars(bonds, 3)
        self.bends, self.bend_par = ints(bends, 4), pars(bends, 4)
        self.torsions, self.torsion_par = ints(torsions, 5), pars(torsions, 5)
        self.intra = ints(intra, 2)
        per = lambda v: _f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)).copy())
        self.max_bond, self.max_bend, self.max_cone = per(max_bond_change), per(max_bend_change), \
            per(max_cone_rotation)
        t = ChainT()
        t.n_atoms, t.start, t.type = n, start, _ptr(self.types)
        t.n_bonds, t.bonds, t.bond_par = len(bonds), _ptr(self.bonds), _ptr(self.bond_par)
        t.n_bends, t.bends, t.bend_par = len(bends), _ptr(self.bends), _ptr(self.bend_par)
        t.n_torsions, t.torsions, t.torsion_par = len(torsions), _ptr(self.torsions), _ptr(self.torsion_par)
        t.n_intra_vdw, t.intra_vdw, t.intra14_vdw_scaling = len(intra), _ptr(self.intra), intra14_scaling
        t.k_fb, t.k_trial, t.k_torsion, t.moves_per_open_bead = k_fb, k_trial, k_torsion, moves_per_open_bead
        t.max_bond_change, t.max_bend_change, t.max_cone_rotation = _ptr(self.max_bond), \
            _ptr(self.max_bend), _ptr(self.max_cone)
        self.struct = t
        self.k_fb, self.k_trial, self.k_torsion = k_fb, k_trial, k_torsion


class WidomBlocks(ctypes.Structure):
    _fields_ = [("count", _D * NR_BLOCKS), ("sum_w", _D * NR_BLOCKS), ("sum_wdu", _D * NR_BLOCKS),
                ("ideal_gas", _D * NR_BLOCKS)]

    def as_array(self):
        """5 x 4 array (count, sum_w, sum_wdu, ideal_gas) per block."""
        return np.array([[self.count[b], self.sum_w[b], self.sum_wdu[b], self.ideal_gas[b]]
                         for b in range(NR_BLOCKS)])


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"librhg.so not built at {LIB_PATH}; run python -m raspa2_amd.build")
    L = ctypes.CDLL(LIB_PATH)

    def f(name, *args, res=ctypes.c_int):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("rhg_version", res=ctypes.c_char_p)
    f("rhg_last_error", _P, res=ctypes.c_char_p)
    f("rhg_create", ctypes.POINTER(Framework), ctypes.POINTER(CellT), ctypes.POINTER(ForceField),
      ctypes.c_int, ctypes.POINTER(_P))
    f("rhg_destroy", _P, res=None)
    f("rhg_add_pseudo_atom", _P, _D, _D, _D, ctypes.POINTER(_I32))
    f("rhg_set_precision", _P, _I32)
    f("rhg_footprint", _P, ctypes.POINTER(ctypes.c_size_t))
    f("rhg_info", _P, _P)
    f("rhg_framework_vdw_energy_at_position", _P, _P, _I32, _D, ctypes.POINTER(_D))
    f("rhg_framework_charge_energy_at_position", _P, _P, _I32, _D, ctypes.POINTER(_D),
      ctypes.POINTER(_D))
    f("rhg_framework_energy_at_positions", _P, _I64, _P, _I32, _P, _P, _P)
    f("rhg_set_tiles", _P, _I32, _I32)
    f("rhg_framework_energy_at_positions_noalloc", _P, _I64, _P, _I32, _P, _P, _P)
    f("rhg_ewald_tables", _P, _P, _P, _P)
    f("rhg_ewald_fourier_adsorbate", _P, _I64, _I32, _P, _P, _P, _P)
    f("rhg_pore_size_distribution", _P, _D, _I64, _I64, _U64, _I32, _D, _P, _P)
    f("rhg_widom_chain_run", _P, ctypes.POINTER(ChainT), _U64, _I64, _I64, _I64, _I32,
      ctypes.POINTER(WidomBlocks), _P, _P)
    f("rhg_widom_run", _P, ctypes.POINTER(MoleculeT), _U64, _I64, _I64, _I64, _I32,
      ctypes.POINTER(WidomBlocks), _P)
    f("rhg_widom_run_trials", _P, ctypes.POINTER(MoleculeT), _I64, _P, _P, _P, _P)
    f("rhg_set_block_pockets", _P, _I32, _P, _P, _P, _I32)
    f("rhg_blocked_pocket", _P, _I64, _P, _P)
    f("rhg_widom_submit", _P, ctypes.POINTER(MoleculeT), _U64, _I64, _I64, _I64, _I32)
    f("rhg_widom_query", _P)
    f("rhg_widom_collect", _P, ctypes.POINTER(WidomBlocks))
    f("rhg_set_widom_options", _P, _I64, _I32)
    f("rhg_inject_fault", _P, _I64)
    f("rhg_widom_estimators", ctypes.POINTER(WidomBlocks), _D, _D, _D, _P)
    f("rhg_make_grid", _P, _I32, _P, _D, _I32, ctypes.POINTER(_P))
    f("rhg_grid_info", _P, _P)
    f("rhg_grid_download", _P, _P, ctypes.c_size_t)
    f("rhg_interpolate_vdw_grid", _P, _I64, _P, _P)
    f("rhg_grid_free", _P, res=None)
    f("rhg_derivatives_at_positions_vdw", _P, _I32, _I64, _P, _P)
    f("rhg_make_coulomb_grid", _P, _P, _D, _I32, ctypes.POINTER(_P))
    f("rhg_interpolate_coulomb_grid", _P, _I64, _P, _D, _P)
    f("rhg_derivatives_at_positions_real", _P, _I64, _P, _P, _P)
    f("rhg_grid_write", _P, ctypes.c_char_p, _P)
    f("rhg_grid_read", _P, ctypes.c_char_p, _I32, _P, _P, ctypes.POINTER(_P))
    f("rhg_grid_meta", _P, _P)
    f("rhg_use_grid", _P, _I32, _P)
    f("rhg_set_adsorbates", _P, _I32, _P, _P, _P)
    f("rhg_inter_energy_at_positions", _P, _I64, _P, _I32, _I32, _P, _P)
    f("rhg_ewald_fourier_move", _P, _I32, _P, _P, _I32, _P)
    f("rhg_widom_retrace_trials", _P, _P, _P, _I64, _P, _P, _P, _P)
    f("rhg_surface_area", _P, _I32, _D, _I64, _U64, _P, _P, _P)
    f("rhg_use_coulomb_grid", _P, _P)
    f("rhg_profile_enable", _P, _I32)
    f("rhg_profile_get", _P, ctypes.c_char_p, ctypes.POINTER(_I64), ctypes.POINTER(_D))
    f("rhg_profile_reset", _P)
    f("rhg_synchronize", _P)
    f("rhg_widom_stats", _P, _P, _I32)
    f("rhg_set_culling", _P, _I32, _D)
    f("rhg_stream_copy", ctypes.c_int, _I64, _I32)
    f("rhg_synthetic_positions", _U64, _P, _I32, _D, _I64, _P, ctypes.POINTER(_I32))
    return L


lib = _load()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _dptr(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def stream_copy(nbytes, reps=3, device=0):
    """Plain device copy kernel of known traffic (counter calibration)."""
    rc = lib.rhg_stream_copy(device, nbytes, reps)
    if rc:
        raise RhgError(f"rhg_stream_copy: status {rc}", rc)


def version():
    return lib.rhg_version().decode()


def molecule(types, body, start=0, k_fb=10, k_or=10):
    m = MoleculeT()
    m.n_sites = len(types)
    for i, t in enumerate(types):
        m.type[i] = int(t)
        for d in range(3):
            m.body[i][d] = float(body[i][d])
    m.start = int(start)
    m.k_fb = int(k_fb)
    m.k_or = int(k_or)
    return m


def widom_estimators(blocks, temperature, density, ideal_gas_weight=1.0):
    out = np.zeros(16)
    rc = lib.rhg_widom_estimators(ctypes.byref(blocks), temperature, density, ideal_gas_weight,
                                  _ptr(out))
    if rc:
        raise RhgError(f"rhg_widom_estimators: {rc}")
    return dict(kh_blocks=out[0:5], kh=out[5], kh_err=out[6], u_blocks=out[7:12], u=out[12],
                u_err=out[13], kh_overall=out[14], w_overall=out[15])


class Grid:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        self._users = weakref.WeakSet()  # contexts using this grid for energies
        ctx._grids.add(self)
        info = np.zeros(13)
        lib.rhg_grid_info(self.h, _ptr(info))
        self.n = info[0:3].astype(int)
        self.size, self.shift, self.delta = info[3:6], info[6:9], info[9:12]
        self.node_bytes = int(info[12])
        self.geom = info[:12].copy()
        meta = np.zeros(10)
        lib.rhg_grid_meta(self.h, _ptr(meta))
        self.spacing = meta[0]
        self.unit_cell_size = meta[1:4].copy()
        self.n_unit_cells = meta[4:7].astype(int)
        self.coulomb = bool(meta[7])
        self.ewald_precision = meta[8]
        self.exec_pairs = meta[9]

    def close(self):
        """Free the device grid now (also done when the object is collected);
        contexts using it for energies stop doing so first."""
        if getattr(self, "h", None):
            for ctx in list(getattr(self, "_users", ())):
                ctx._detach(self)
            lib.rhg_grid_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self.n[0] + 1, self.n[1] + 1, self.n[2] + 1, 8)

    def download(self):
        dt = np.float64 if self.node_bytes == 64 else np.float32
        out = np.zeros(self.shape, dtype=dt)
        self.ctx._check(lib.rhg_grid_download(self.h, _ptr(out), out.nbytes), "grid_download")
        return out

    def interpolate(self, xyz, charge=None):
        """InterpolateVDWGrid (src/grids.c:979) at many positions; for a
        Coulomb grid InterpolateCoulombGrid (src/grids.c:1200) of a probe of
        the given charge."""
        xyz = _f64(xyz).reshape(-1, 3)
        u = np.zeros(len(xyz))
        if self.coulomb:
            if charge is None:
                raise RhgError("a Coulomb grid is interpolated for a probe charge", ERR_INVALID)
            self.ctx._check(lib.rhg_interpolate_coulomb_grid(self.h, len(xyz), _ptr(xyz), float(charge),
                                                             _ptr(u)), "interpolate_coulomb_grid")
        else:
            self.ctx._check(lib.rhg_interpolate_vdw_grid(self.h, len(xyz), _ptr(xyz), _ptr(u)),
                            "interpolate_vdw_grid")
        return u

    def write(self, path, n_unit_cells=None):
        """WriteVDWGrid / WriteCoulombGrid (src/grids.c:692, 815) to `path`."""
        nc = None if n_unit_cells is None else np.ascontiguousarray(n_unit_cells, dtype=np.int32)
        self.ctx._check(lib.rhg_grid_write(self.h, os.fsencode(path), None if nc is None else _ptr(nc)),
                        "grid_write")


def grid_path(raspa_dir, forcefield, framework, spacing, pseudo_atom=None, shifted=True,
              n_unit_cells=None, method="Ewald"):
    """Where RASPA2 keeps a grid (src/grids.c:706-724, 838-856): the VDW grid
    of `pseudo_atom`, or with pseudo_atom None the Coulomb grid of a system of
    n_unit_cells unit cells."""
    base = os.path.join(raspa_dir, "share", "raspa", "grids", forcefield, framework, "%f" % spacing)
    if pseudo_atom is not None:
        return os.path.join(base, "%s_%s_%s.grid" % (framework, pseudo_atom,
                                                     "shifted" if shifted else "truncated"))
    nc = "%dx%dx%d" % tuple(int(v) for v in n_unit_cells)
    return os.path.join(base, nc, "%s_Electrostatics_%s.grid" % (framework, method))


class Context:
    """One framework resident on one GPU (rhg_ctx)."""

    def __init__(self, box, bc, x, y, z, eps_k, sigma, charge=None, mass=None, cutoff_vdw=12.0,
                 shifted=True, charge_method=CHARGE_NONE, cutoff_charge=12.0,
                 ewald_precision=1e-6, temperature=300.0, device=0, precision=PREC_FP64):
        self._arrs = [_f64(x), _f64(y), _f64(z), _f64(eps_k), _f64(sigma),
                      _f64(charge) if charge is not None else None,
                      _f64(mass) if mass is not None else None]
        xx, yy, zz, ee, ss, qq, mm = self._arrs
        fw = Framework(len(xx), _dptr(xx), _dptr(yy), _dptr(zz), _dptr(qq), _dptr(ee), _dptr(ss),
                       _dptr(mm))
        cell = CellT()
        for i, v in enumerate(np.asarray(box, dtype=np.float64).reshape(9)):
            cell.box[i] = v
        cell.boundary = bc
        ff = ForceField(cutoff_vdw, 1 if shifted else 0, charge_method, cutoff_charge,
                        ewald_precision, temperature)
        h = _P()
        self._grids = weakref.WeakSet()
        self._in_use = {}
        rc = lib.rhg_create(ctypes.byref(fw), ctypes.byref(cell), ctypes.byref(ff), device,
                            ctypes.byref(h))
        if rc:
            raise RhgError(f"rhg_create failed: {rc}")
        self.h = h
        self.temperature = temperature
        self.n_atoms = len(xx)
        self.set_precision(precision)

    def close(self):
        # grids run on this context's stream: they go first, also when the
        # garbage collector finalises a cycle in arbitrary order
        for g in list(getattr(self, "_grids", ())):
            g.close()
        if getattr(self, "h", None):
            lib.rhg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = lib.rhg_last_error(self.h)
            raise RhgError(f"{what}: status {rc}: {msg.decode() if msg else ''}", rc)

    def add_pseudo_atom(self, eps_k, sigma, charge=0.0):
        t = _I32()
        self._check(lib.rhg_add_pseudo_atom(self.h, eps_k, sigma, charge, ctypes.byref(t)),
                    "add_pseudo_atom")
        return t.value

    def set_precision(self, p):
        self._check(lib.rhg_set_precision(self.h, p), "set_precision")

    def info(self):
        out = np.zeros(10)
        self._check(lib.rhg_info(self.h, _ptr(out)), "info")
        return dict(alpha=out[0], kx=int(out[1]), ky=int(out[2]), kz=int(out[3]), nk=int(out[4]),
                    uion=out[5], beta=out[6], volume=out[7], mass=out[8], density=out[9])

    def footprint(self):
        b = ctypes.c_size_t()
        self._check(lib.rhg_footprint(self.h, ctypes.byref(b)), "footprint")
        return b.value

    # ---- src/framework_energy.c:60,176
    def CalculateFrameworkVDWEnergyAtPosition(self, pos, type_a, scaling=1.0):
        p = _f64(pos)
        u = _D()
        self._check(lib.rhg_framework_vdw_energy_at_position(self.h, _ptr(p), type_a, scaling,
                                                               ctypes.byref(u)), "vdw_at")
        return u.value

    def CalculateFrameworkChargeEnergyAtPosition(self, pos, type_a, scaling=1.0):
        p = _f64(pos)
        ucc, ucbd = _D(), _D()
        self._check(lib.rhg_framework_charge_energy_at_position(
            self.h, _ptr(p), type_a, scaling, ctypes.byref(ucc), ctypes.byref(ucbd)), "charge_at")
        return ucc.value, ucbd.value

    def energy_at_positions(self, xyz, type_a):
        xyz = _f64(xyz).reshape(-1, 3)
        n = len(xyz)
        uv, uc = np.zeros(n), np.zeros(n)
        ov = np.zeros(n, dtype=np.uint8)
        self._check(lib.rhg_framework_energy_at_positions(self.h, n, _ptr(xyz), type_a, _ptr(uv),
                                                          _ptr(uc), _ptr(ov)), "energy_at_positions")
        return uv, uc, ov.astype(bool)

    # ---- src/ewald.c
    def ewald_tables(self):
        nk = self.info()["nk"]
        kf, sr, si = np.zeros(nk), np.zeros(nk), np.zeros(nk)
        self._check(lib.rhg_ewald_tables(self.h, _ptr(kf), _ptr(sr), _ptr(si)), "ewald_tables")
        return dict(kfactor=kf, sf_re=sr, sf_im=si)

    def CalculateEwaldFourierAdsorbate(self, types, xyz):
        """xyz: n_mol x n_sites x 3; returns (d_host, d_ads) arrays."""
        types = np.ascontiguousarray(types, dtype=np.int32)
        xyz = _f64(xyz).reshape(-1, len(types), 3)
        n = len(xyz)
        dh, da = np.zeros(n), np.zeros(n)
        self._check(lib.rhg_ewald_fourier_adsorbate(self.h, n, len(types), _ptr(types), _ptr(xyz),
                                                    _ptr(dh), _ptr(da)), "fourier_adsorbate")
        return dh, da

    # ---- src/mc_moves.c:7844 WidomAdsorbateMove
    def widom_run(self, mol, seed, n_total, first=0, n=None, steps_per_cycle=20, blocks=None,
                  records=False):
        n = n_total - first if n is None else n
        blocks = WidomBlocks() if blocks is None else blocks
        rec = np.zeros((n, REC_N)) if records else None
        self._check(lib.rhg_widom_run(self.h, ctypes.byref(mol), seed, n_total, first, n,
                                      steps_per_cycle, ctypes.byref(blocks), _ptr(rec)),
                    "widom_run")
        return blocks, rec

    def widom_chain_run(self, chain, seed, n_total, first=0, n=None, steps_per_cycle=20, blocks=None,
                        records=False, positions=False):
        """WidomAdsorbateMove of a flexible chain (GrowMolecule with the
        internal MC scheme, src/cbmc.c:3056-3185, 1580-2482) for insertions
        [first, first+n); returns (blocks, records or None, positions or None)."""
        n = n_total - first if n is None else n
        blocks = WidomBlocks() if blocks is None else blocks
        rec = np.zeros((n, CREC_N)) if records else None
        xyz = np.zeros((n, chain.n_atoms, 3)) if positions else None
        self._check(lib.rhg_widom_chain_run(self.h, ctypes.byref(chain.struct), seed, n_total, first, n,
                                            steps_per_cycle, ctypes.byref(blocks), _ptr(rec), _ptr(xyz)),
                    "widom_chain_run")
        return blocks, rec, xyz

    def PoreSizeDistribution(self, cycles, n_b=None, seed=1, bins=50, range_=12.5, probe_factor=1.0,
                             largest=False):
        """SamplePoreSizeDistribution over `cycles` points with n_b sphere
        centres each (default 20 x cycles, the reference's count): the
        cumulative histogram PoreSizeDistributionHistogram[bins] and, when
        asked, each cycle's largest radius."""
        n_b = 20 * cycles if n_b is None else n_b
        hist = np.zeros(bins, dtype=np.uint64)
        lg = np.zeros(cycles) if largest else None
        self._check(lib.rhg_pore_size_distribution(self.h, probe_factor, cycles, n_b, seed, bins, range_,
                                                   _ptr(hist), _ptr(lg)), "pore_size_distribution")
        return (hist, lg) if largest else hist

    # ---- src/framework.c:6168-6300 blocked pockets
    def set_block_pockets(self, centres, radii, unit_box, invert=False):
        c = _f64(centres).reshape(-1, 3)
        r = _f64(radii).reshape(-1)
        ub = _f64(unit_box).reshape(9)
        self._check(lib.rhg_set_block_pockets(self.h, len(r), _ptr(c), _ptr(r), _ptr(ub),
                                              1 if invert else 0), "set_block_pockets")

    def BlockedPocket(self, xyz):
        xyz = _f64(xyz).reshape(-1, 3)
        out = np.zeros(len(xyz), dtype=np.uint8)
        self._check(lib.rhg_blocked_pocket(self.h, len(xyz), _ptr(xyz), _ptr(out)),
                    "blocked_pocket")
        return out.astype(bool)

    def widom_submit(self, mol, seed, n_total, first=0, n=None, steps_per_cycle=20):
        """Enqueue insertions and return at once (sums accumulate on the device)."""
        n = n_total - first if n is None else n
        self._check(lib.rhg_widom_submit(self.h, ctypes.byref(mol), seed, n_total, first, n,
                                         steps_per_cycle), "widom_submit")

    def widom_query(self):
        """True when every submitted insertion has finished."""
        rc = lib.rhg_widom_query(self.h)
        if rc < 0:
            self._check(rc, "widom_query")
        return rc == 1

    def widom_collect(self, blocks=None):
        """Wait, add the accumulated block sums to `blocks` and clear them."""
        blocks = WidomBlocks() if blocks is None else blocks
        self._check(lib.rhg_widom_collect(self.h, ctypes.byref(blocks)), "widom_collect")
        return blocks

    def set_widom_options(self, chunk=1 << 20, graphs=True):
        self._check(lib.rhg_set_widom_options(self.h, chunk, 1 if graphs else 0),
                    "set_widom_options")

    def inject_fault(self, after_chunks):
        self._check(lib.rhg_inject_fault(self.h, after_chunks), "inject_fault")

    def widom_run_trials(self, mol, fb, cord, usel):
        fb = _f64(fb)
        n = fb.shape[0]
        cord = _f64(cord) if cord is not None else None
        usel = _f64(usel)
        rec = np.zeros((n, REC_N))
        self._check(lib.rhg_widom_run_trials(self.h, ctypes.byref(mol), n, _ptr(fb), _ptr(cord),
                                             _ptr(usel), _ptr(rec)), "widom_run_trials")
        return rec

    # ---- src/grids.c
    def MakeGrid(self, type_a, unit_box, spacing, fp64=False):
        ub = _f64(unit_box).reshape(9)
        h = _P()
        self._check(lib.rhg_make_grid(self.h, type_a, _ptr(ub), spacing, 1 if fp64 else 0,
                                      ctypes.byref(h)), "make_grid")
        return Grid(self, h)

    def MakeCoulombGrid(self, unit_box, spacing, fp64=False):
        """MakeGrid's Ewald-Coulomb grid (src/grids.c:587-689)."""
        ub = _f64(unit_box).reshape(9)
        h = _P()
        self._check(lib.rhg_make_coulomb_grid(self.h, _ptr(ub), spacing, 1 if fp64 else 0,
                                              ctypes.byref(h)), "make_coulomb_grid")
        return Grid(self, h)

    def ReadGrid(self, path, unit_box, coulomb=False, n_unit_cells=None):
        """ReadVDWGrid / ReadCoulombGrid (src/grids.c:745, 879) from `path`."""
        ub = _f64(unit_box).reshape(9)
        nc = None if n_unit_cells is None else np.ascontiguousarray(n_unit_cells, dtype=np.int32)
        h = _P()
        self._check(lib.rhg_grid_read(self.h, os.fsencode(path), 1 if coulomb else 0, _ptr(ub),
                                      None if nc is None else _ptr(nc), ctypes.byref(h)), "grid_read")
        return Grid(self, h)

    # ---- finite loading (src/inter_energy.c, src/ewald.c:6859)
    def set_adsorbates(self, start, types, xyz):
        """The adsorbate molecules present: molecule m holds atoms
        [start[m], start[m+1]) of the given pseudo-atom types at xyz."""
        st = np.ascontiguousarray(start, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        xy = _f64(xyz).reshape(-1, 3)
        n_mol = max(len(st) - 1, 0)
        self._check(lib.rhg_set_adsorbates(self.h, n_mol, _ptr(st) if n_mol else None,
                                           _ptr(ty) if n_mol else None, _ptr(xy) if n_mol else None),
                    "set_adsorbates")

    def CalculateInterEnergyAdsorbateAtPosition(self, xyz, type_a, exclude=-1):
        """(u_vdw, u_cc) of pseudo-atom type_a at n positions against every
        adsorbate molecule but `exclude`."""
        xyz = _f64(xyz).reshape(-1, 3)
        uv, uc = np.zeros(len(xyz)), np.zeros(len(xyz))
        self._check(lib.rhg_inter_energy_at_positions(self.h, len(xyz), _ptr(xyz), type_a, exclude,
                                                      _ptr(uv), _ptr(uc)), "inter_energy")
        return uv, uc

    def RetraceMolecule(self, mol, exclude, old_xyz, fb, cord):
        """Rosenbluth weights of existing molecules (CBMC_DELETION): old_xyz
        n x n_sites x 3, fb n x k_fb x 3, cord n x k_or x n_sites x 3 (entries
        0 of the trials are taken by the molecules' own configuration)."""
        old_xyz = _f64(old_xyz)
        n = old_xyz.shape[0]
        fb = _f64(fb)
        cord = _f64(cord)
        ex = np.ascontiguousarray(np.broadcast_to(np.asarray(exclude, dtype=np.int32), (n,)))
        rec = np.zeros((n, REC_N))
        self._check(lib.rhg_widom_retrace_trials(self.h, ctypes.byref(mol), _ptr(ex), n, _ptr(old_xyz),
                                                 _ptr(fb), _ptr(cord), _ptr(rec)), "retrace")
        return rec

    def SurfaceArea(self, probe_type, samples, seed=1, distance_factor=2.0 ** (1.0 / 6.0), interacts=None):
        """(area [A^2], accessible points per atom) of SurfaceAreaMove with
        `samples` points per framework atom; interacts (optional, per atom)
        False for atoms of NONE pseudo-atoms."""
        area = _D()
        counts = np.zeros(self.n_atoms, dtype=np.uint64)
        mask = None if interacts is None else np.ascontiguousarray(interacts, dtype=np.uint8)
        self._check(lib.rhg_surface_area(self.h, probe_type, distance_factor, samples, seed,
                                         None if mask is None else _ptr(mask),
                                         ctypes.byref(area), _ptr(counts)), "surface_area")
        return area.value, counts

    def CalculateEwaldFourierMove(self, new_types=(), new_xyz=None, old_mol=-1):
        """(host, adsorbate) Fourier changes of inserting a molecule and/or
        removing adsorbate old_mol."""
        ty = np.ascontiguousarray(list(new_types), dtype=np.int32)
        xy = _f64(new_xyz if new_xyz is not None else np.zeros((0, 3))).reshape(-1, 3)
        out = np.zeros(2)
        self._check(lib.rhg_ewald_fourier_move(self.h, len(ty), _ptr(ty) if len(ty) else None,
                                               _ptr(xy) if len(ty) else None, old_mol, _ptr(out)),
                    "fourier_move")
        return out

    def use_grid(self, type_a, grid):
        """UseTabularGrid for pseudo-atom type_a (grid None detaches); the
        context keeps the grid alive while it is in use."""
        self._check(lib.rhg_use_grid(self.h, type_a, grid.h if grid is not None else None), "use_grid")
        self._in_use[type_a] = grid
        if grid is not None:
            grid._users.add(self)

    def use_coulomb_grid(self, grid):
        self._check(lib.rhg_use_coulomb_grid(self.h, grid.h if grid is not None else None),
                    "use_coulomb_grid")
        self._in_use["coulomb"] = grid
        if grid is not None:
            grid._users.add(self)

    def _detach(self, grid):
        if not getattr(self, "h", None):
            return
        for key, g in list(self._in_use.items()):
            if g is grid:
                if key == "coulomb":
                    lib.rhg_use_coulomb_grid(self.h, None)
                else:
                    lib.rhg_use_grid(self.h, key, None)
                self._in_use[key] = None

    def CalculateDerivativesAtPositionReal(self, xyz):
        """(out n x 14, smallest_r n) of src/framework_hessian.c:111."""
        xyz = _f64(xyz).reshape(-1, 3)
        out = np.zeros((len(xyz), 14))
        rmin = np.zeros(len(xyz))
        self._check(lib.rhg_derivatives_at_positions_real(self.h, len(xyz), _ptr(xyz), _ptr(out),
                                                          _ptr(rmin)), "derivatives_real")
        return out, rmin

    def CalculateDerivativesAtPositionVDW(self, xyz, type_a):
        xyz = _f64(xyz).reshape(-1, 3)
        out = np.zeros((len(xyz), 14))
        self._check(lib.rhg_derivatives_at_positions_vdw(self.h, type_a, len(xyz), _ptr(xyz),
                                                         _ptr(out)), "derivatives")
        return out

    # ---- profiling
    def profile(self, on=True, site_energy_only=False):
        """HIP-event timing of the kernel families; site_energy_only times the
        site-energy kernels alone (RHG_PROFILE_SITE_ENERGY)."""
        mode = (2 if site_energy_only else 1) if on else 0
        self._check(lib.rhg_profile_enable(self.h, mode), "profile_enable")

    def profile_get(self, name):
        n, ms = _I64(), _D()
        self._check(lib.rhg_profile_get(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)),
                    "profile_get")
        return n.value, ms.value

    def profile_reset(self):
        self._check(lib.rhg_profile_reset(self.h), "profile_reset")

    def set_culling(self, on=True, cell_size=2.5, images=True):
        """Cell-list culling; `images` also enables the image lists of
        the mixed-precision kernels."""
        mode = 0 if not on else (2 if images else 1)
        self._check(lib.rhg_set_culling(self.h, mode, cell_size), "set_culling")

    def set_tiles(self, sites_per_lane=4, waves_per_block=4):
        """Launch shape of the culled site kernels (results are unchanged)."""
        self._check(lib.rhg_set_tiles(self.h, sites_per_lane, waves_per_block), "set_tiles")

    def widom_stats(self, reset=False):
        out = np.zeros(12)
        self._check(lib.rhg_widom_stats(self.h, _ptr(out), 1 if reset else 0), "widom_stats")
        keys = ("insertions", "fb_sites", "or_sites", "cc_sites", "fourier_mols", "lj_pairs",
                "ljcc_pairs", "cc_pairs", "k_site_terms", "exec_lj_pairs", "exec_ljcc_pairs",
                "exec_cc_pairs")
        return dict(zip(keys, out))

    def synchronize(self):
        self._check(lib.rhg_synchronize(self.h), "synchronize")
