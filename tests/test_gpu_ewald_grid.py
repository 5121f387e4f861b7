"""Ewald tables, Fourier Delta and energy grids on the GPU against the oracle
(src/ewald.c:200-1731,6859-7565; src/grids.c:431-1077;
src/framework_hessian.c:63-109)."""
import numpy as np
import pytest

from helpers import oracle_system, random_positions, same_bits

pytestmark = pytest.mark.gpu


def test_ewald_tables(W, O, R):
    fw = W.c1_framework()
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    io, ig = sys_o.info(), ctx.info()
    assert (io["kx"], io["ky"], io["kz"], io["nk"]) == (ig["kx"], ig["ky"], ig["kz"], ig["nk"])
    assert io["alpha"] == ig["alpha"] and io["uion"] == ig["uion"]
    to, tg = sys_o.ewald_tables(), ctx.ewald_tables()
    assert same_bits(to["kfactor"], tg["kfactor"])
    s_o = to["sf_re"] + 1j * to["sf_im"]
    s_g = tg["sf_re"] + 1j * tg["sf_im"]
    scale = np.sqrt(np.sum(np.abs(fw.charge) ** 2))
    assert np.max(np.abs(s_o - s_g)) <= 1e-12 * scale * np.sqrt(fw.n)


def test_fourier_adsorbate(W, O, R):
    fw = W.c1_framework()
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    to = [sys_o.add_pseudo_atom(*s) for s in W.CO2.sites]
    tg = [ctx.add_pseudo_atom(*s) for s in W.CO2.sites]
    centers = random_positions(fw.box, 200, 21)
    rng = np.random.default_rng(3)
    xyz = np.zeros((200, 3, 3))
    for i, c in enumerate(centers):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        xyz[i] = [c + 1.16 * v, c, c - 1.16 * v]
    dh, da = ctx.CalculateEwaldFourierAdsorbate(tg, xyz)
    for i in range(200):
        out = sys_o.fourier_adsorbate(to, xyz[i])
        assert abs(dh[i] - out[0]) <= 1e-10 * max(1.0, abs(out[0]))
        assert abs(da[i] - out[1]) <= 1e-10 * max(1.0, abs(out[1]))


@pytest.mark.parametrize("cull", [False, True])
def test_grid_nodes_bit_exact(W, O, R, cull):
    """MakeGrid node values (FP32 channels) equal the oracle's bit for bit on a
    coarse grid of the triclinic C2 framework, built all-pairs or through the
    cell lists."""
    fw = W.c2_framework()
    sys_o = oracle_system(fw)
    ctx = fw.context()
    ctx.set_culling(cull, 2.5)
    to = sys_o.add_pseudo_atom(148.0, 3.73, 0.0)
    tg = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
    g = ctx.MakeGrid(tg, fw.unit_box, 1.0)
    geom = O.grid_geometry(fw.unit_box, 1.0)
    assert np.array_equal(g.geom, geom)
    data = g.download()
    rng = np.random.default_rng(5)
    idx = [(0, 0, 0), tuple(np.array(data.shape[:3]) - 1)]
    idx += [tuple(rng.integers(0, s) for s in data.shape[:3]) for _ in range(300)]
    for (i, j, k) in idx:
        ref = sys_o.grid_node(to, geom, int(i), int(j), int(k))
        assert np.array_equal(data[i, j, k].view(np.uint32), ref.view(np.uint32)), (i, j, k)


def test_grid_interpolation(W, O, R):
    """InterpolateVDWGrid: tensor-product Hermite on the device equals the
    Lekien-Marsden matrix form of the oracle on the same grid."""
    fw = W.c2_framework()
    ctx = fw.context()
    tg = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
    g = ctx.MakeGrid(tg, fw.unit_box, 0.5)
    data = g.download()
    geom = O.grid_geometry(fw.unit_box, 0.5)
    inv, _ = O.invert3x3(fw.unit_box)
    pts = random_positions(fw.box, 2000, 31)
    u = g.interpolate(pts)
    for p, ug in zip(pts[:400], u[:400]):
        ref = O.interpolate_grid(data, geom, O.BC_TRICLINIC, fw.unit_box, inv, p)
        assert abs(ug - ref) <= 1e-9 * max(1.0, abs(ref)), (p, ug, ref)


def test_derivatives(W, O, R):
    fw = W.c2_framework()
    sys_o = oracle_system(fw)
    ctx = fw.context()
    to = sys_o.add_pseudo_atom(148.0, 3.73, 0.0)
    tg = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
    pts = random_positions(fw.box, 500, 41)
    d = ctx.CalculateDerivativesAtPositionVDW(pts, tg)
    for i in range(500):
        ref = sys_o.derivatives_vdw_at(to, pts[i])
        assert same_bits(d[i], ref), i


def test_grid_mixed_vs_fp64(W, O, R):
    """Mixed-precision grid build (FP32 pair terms over the fixed-point image
    lists, FP64 sums): the same capped nodes as the FP64 build, and the
    interpolated energy within 1e-5 relative (|U| > 1 K) or 1e-3 K."""
    fw = W.c2_framework()
    grids = []
    for prec in (R.PREC_FP64, R.PREC_MIXED):
        ctx = fw.context(precision=prec)
        tg = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
        grids.append((ctx, ctx.MakeGrid(tg, fw.unit_box, 0.5)))
    d64, dmx = grids[0][1].download(), grids[1][1].download()
    assert np.array_equal(d64[..., 0] >= 2e7, dmx[..., 0] >= 2e7)
    pts = random_positions(fw.box, 4000, 37)
    u64, umx = grids[0][1].interpolate(pts), grids[1][1].interpolate(pts)
    e2k = O.units()["energy_to_kelvin"]
    ok = np.abs(u64) < 1e7
    err = np.abs(umx[ok] - u64[ok]) * e2k
    big = np.abs(u64[ok]) * e2k > 1.0
    assert np.all(err[big] <= 1e-5 * np.abs(u64[ok][big]) * e2k)
    assert np.all(err[~big] <= 1e-3)
