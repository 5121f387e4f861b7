"""rhg_framework_energy_at_positions_noalloc (no counterpart in the
reference; for callers that cannot allocate device memory) against the
allocating rhg_framework_energy_at_positions on the same positions: both run
the same kernels, so energies and overlap flags are bit-identical.  Two calls
in a row, the second with fewer positions, share the library's fixed device
array; the size limit and the missing-cell-list case are refused on the host."""
import ctypes

import numpy as np
import pytest

from helpers import DeviceArray, random_positions, same_bits

pytestmark = pytest.mark.gpu

MAX_SITES = 1 << 18


def _device_call(R, ctx, xyz, t, with_cc=True):
    """The no-alloc entry point on device buffers: returns (rc, u_vdw, u_cc, overlap)."""
    n = len(xyz)
    d_xyz = DeviceArray(np.ascontiguousarray(xyz, dtype=np.float64))
    d_uv = DeviceArray(np.full(n, -1.0))
    d_uc = DeviceArray(np.full(n, -1.0))
    d_ov = DeviceArray(np.full(n, 7, dtype=np.uint8))
    try:
        rc = R.lib.rhg_framework_energy_at_positions_noalloc(
            ctx.h, n, d_xyz.ptr, t, d_uv.ptr, d_uc.ptr if with_cc else None, d_ov.ptr)
        ctx.synchronize()
        return rc, d_uv.get(), d_uc.get(), d_ov.get().astype(bool)
    finally:
        for d in (d_xyz, d_uv, d_uc, d_ov):
            d.free()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_allocating_form(W, R, precision):
    print(f"noalloc vs allocating form: C1 supercell, CO2 oxygen probe, {precision}", flush=True)
    fw = W.c1_framework()
    prec = R.PREC_FP64 if precision == "fp64" else R.PREC_MIXED
    ctx = fw.context(charge_method=R.CHARGE_EWALD, precision=prec)
    t = ctx.add_pseudo_atom(79.0, 3.05, -0.35)
    for n, seed in ((200000, 11), (3001, 12)):
        xyz = random_positions(fw.box, n, seed)
        uv, uc, ov = ctx.energy_at_positions(xyz, t)  # also builds the cell and image lists
        rc, duv, duc, dov = _device_call(R, ctx, xyz, t)
        assert rc == 0, R.lib.rhg_last_error(ctx.h)
        assert same_bits(duv, uv) and same_bits(duc, uc)
        assert np.array_equal(dov, ov)
        assert 0 < ov.sum() < n
    ctx.close()


def test_noalloc_uncharged_probe_and_null_output(W, R):
    """An uncharged probe gets zero Coulomb energies; a null output is
    skipped (C0 cell, methane, FP64)."""
    fw = W.c0_framework()
    ctx = fw.context(precision=R.PREC_FP64)
    t = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
    xyz = random_positions(fw.box, 5000, 3)
    uv, uc, ov = ctx.energy_at_positions(xyz, t)
    rc, duv, duc, dov = _device_call(R, ctx, xyz, t)
    assert rc == 0
    assert same_bits(duv, uv) and np.all(duc == 0.0) and np.array_equal(dov, ov)
    rc, duv2, _, _ = _device_call(R, ctx, xyz[:777], t, with_cc=False)
    assert rc == 0 and same_bits(duv2, uv[:777])
    ctx.close()


def test_noalloc_refusals(W, R):
    """Sizes beyond the fixed array and contexts whose cell lists are not yet
    built are refused before anything is launched."""
    fw = W.c0_framework()
    ctx = fw.context(precision=R.PREC_MIXED)
    t = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
    xyz = random_positions(fw.box, 10, 4)
    rc, _, _, _ = _device_call(R, ctx, xyz, t)
    assert rc == R.ERR_INVALID and b"cell lists" in R.lib.rhg_last_error(ctx.h)
    ctx.energy_at_positions(xyz, t)
    rc = R.lib.rhg_framework_energy_at_positions_noalloc(ctx.h, MAX_SITES + 1, ctypes.c_void_p(16), t,
                                                        None, None, None)
    assert rc == R.ERR_INVALID and b"262144" in R.lib.rhg_last_error(ctx.h)
    rc, _, _, _ = _device_call(R, ctx, xyz, t)
    assert rc == 0
    ctx.close()
