"""Builders shared by the tests: the same framework on the oracle and on the GPU."""
import numpy as np

import oracle as O
from raspa2_amd import rhg as R
from raspa2_amd import workloads as W


def oracle_system(fw, charge_method=O.CHARGE_NONE, cutoff=12.0, precision=1e-6, T=300.0):
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    return O.System(box, bc, x, y, z, q, eps, sig, mass=mass, cutoff_vdw=cutoff, shifted=True,
                    charge_method=charge_method, cutoff_charge=cutoff,
                    ewald_precision=precision, temperature=T)


def oracle_molecule(sys_o, probe, k_fb=None, k_or=None):
    types = [sys_o.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    return types, O.molecule(types, probe.body, start=probe.start,
                             k_fb=probe.k_fb if k_fb is None else k_fb,
                             k_or=probe.k_or if k_or is None else k_or)


def gpu_molecule(ctx, probe, k_fb=None, k_or=None):
    types = [ctx.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    return types, R.molecule(types, probe.body, start=probe.start,
                             k_fb=probe.k_fb if k_fb is None else k_fb,
                             k_or=probe.k_or if k_or is None else k_or)


def oracle_chain(sys_o, probe, **kw):
    """The chain probe on the oracle system (same pseudo-atom order as the GPU)."""
    types = [sys_o.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    args = dict(start=probe.start, k_fb=probe.k_fb, k_trial=probe.k_trial, k_torsion=probe.k_torsion,
                moves=probe.moves)
    args.update(kw)
    return types, O.chain(types, probe.bonds, probe.bends, probe.torsions, probe.intra, **args)


def gpu_chain(ctx, probe, **kw):
    types = [ctx.add_pseudo_atom(e, s, q) for e, s, q in probe.sites]
    args = dict(start=probe.start, k_fb=probe.k_fb, k_trial=probe.k_trial, k_torsion=probe.k_torsion,
                moves_per_open_bead=probe.moves)
    args.update(kw)
    return types, R.Chain(types, probe.bonds, probe.bends, probe.torsions, probe.intra, **args)


def random_positions(box, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = rng.uniform(0.0, 1.0, (n, 3))
    m = np.asarray(box).reshape(3, 3)  # rows a, b, c
    return s @ m


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def irmof1_dubbeldam():
    """The reference's grid example decks (IRMOF-1, Dubbeldam2007FlexibleIRMOF-1,
    TraPPE CO2) from the committed fixtures."""
    import os
    return W.irmof1_dubbeldam(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))


class DeviceArray:
    """A device allocation through the HIP runtime librhg links against (the
    test's own memory; used to hand device pointers to the C ABI)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        import ctypes
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        if self.host.nbytes:
            assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.host)
        if out.nbytes:
            assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
