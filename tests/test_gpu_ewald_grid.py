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
    """Mixed-precision grid build (FP32 pair terms over the image
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


# ---------------------------------------------------------------- Coulomb grid
def _charged(W, name):
    """c1: the 25.8 A cubic cell replicated 2x2x2 (rectangular); c2q: the
    triclinic C2 cell with C1-style charges."""
    if name == "c1":
        return W.c1_framework()
    fw = W.c2_framework()
    fw.charge = W._charges(1202, fw.n)
    return fw


def _f32_close(a, b, atol):
    """Channels within two float32 ulps of the oracle's, plus `atol` for the
    FP64 sums (device erfc/exp against libm's)."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    tol = 2.0 * np.spacing(np.abs(b)) + atol
    return np.all(np.abs(a.astype(np.float64) - b.astype(np.float64)) <= tol)


def test_real_derivatives(W, O, R):
    """CalculateDerivativesAtPositionReal on the device: the smallest distance
    bit for bit, the 14 values within 1e-12 of each column's scale."""
    fw = W.c1_framework()
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    pts = random_positions(fw.box, 400, 43)
    d, rmin = ctx.CalculateDerivativesAtPositionReal(pts)
    ref = np.zeros_like(d)
    rref = np.zeros(len(pts))
    for i, p in enumerate(pts):
        ref[i], rref[i] = sys_o.derivatives_real_at(p)
    assert same_bits(rmin, rref)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(d - ref) <= 1e-12 * scale), np.abs(d - ref).max(axis=0) / scale


@pytest.mark.parametrize("name", ["c1", "c2q"])
@pytest.mark.parametrize("cull", [False, True])
def test_coulomb_grid_nodes(W, O, R, name, cull):
    """Ewald-Coulomb grid nodes against the oracle's: the same capped nodes,
    FP32 channels within two ulps (plus 1e-9 of the channel's scale for the
    device's erfc and exp)."""
    fw = _charged(W, name)
    sys_o = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    ctx.set_culling(cull, 2.5)
    g = ctx.MakeCoulombGrid(fw.unit_box, 1.0)
    geom = O.grid_geometry(fw.unit_box, 1.0)
    assert np.array_equal(g.geom, geom) and g.coulomb
    data = g.download()
    rng = np.random.default_rng(6)
    idx = [(0, 0, 0), tuple(np.array(data.shape[:3]) - 1)]
    idx += [tuple(rng.integers(0, s) for s in data.shape[:3]) for _ in range(250)]
    refs = np.array([sys_o.coulomb_grid_node(geom, int(i), int(j), int(k)) for i, j, k in idx])
    got = np.array([data[i, j, k] for i, j, k in idx])
    capped = refs[:, 0] == np.float32(2e7)
    assert np.array_equal(got[:, 0] == np.float32(2e7), capped)
    live = refs[~capped]
    atol = 1e-9 * np.abs(live).max(axis=0)
    for ch in range(8):
        assert _f32_close(got[~capped, ch], live[:, ch], atol[ch]), ch


def test_coulomb_grid_interpolation(W, O, R):
    """InterpolateCoulombGrid on the device equals the oracle's Lekien-Marsden
    form on the same grid, times the probe charge; on a 0.2 A grid it tracks
    the direct real-space energy away from the atoms."""
    fw = W.c1_framework()
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    g = ctx.MakeCoulombGrid(fw.unit_box, 0.5)
    data = g.download()
    geom = O.grid_geometry(fw.unit_box, 0.5)
    inv, _ = O.invert3x3(fw.unit_box)
    pts = random_positions(fw.box, 2000, 47)
    u = g.interpolate(pts, charge=-0.35)
    for p, ug in zip(pts[:300], u[:300]):
        ref = O.interpolate_coulomb_grid(data, geom, O.BC_RECTANGULAR, fw.unit_box, inv, p, -0.35)
        assert abs(ug - ref) <= 1e-9 * max(1.0, abs(ref)), (p, ug, ref)
    with pytest.raises(R.RhgError):
        g.interpolate(pts)
    g.close()
    fine = ctx.MakeCoulombGrid(fw.unit_box, 0.2)
    t = ctx.add_pseudo_atom(0.0, 1.0, -0.35)
    _, rmin = ctx.CalculateDerivativesAtPositionReal(pts)
    far = rmin > 2.0
    _, uc, _ = ctx.energy_at_positions(pts[far], t)
    ui = fine.interpolate(pts[far], charge=-0.35)
    rel = np.abs(ui - uc) / np.maximum(np.abs(uc), 1.0)
    assert np.median(rel) < 1e-4 and rel.max() < 1e-2, (np.median(rel), rel.max())


# ---------------------------------------------------------------- grid files
_HDR = "<d3i3d3d3d3d3i"


def _parse(path, coulomb):
    import struct
    raw = open(path, "rb").read()
    n = struct.calcsize(_HDR)
    assert n == 128
    h = struct.unpack_from(_HDR, raw, 0)
    off = n
    ep = None
    if coulomb:
        ep = struct.unpack_from("<d", raw, n)[0]
        off += 8
    N = h[1:4]
    nodes = (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
    assert len(raw) == off + nodes * 8 * 4
    data = np.frombuffer(raw, dtype="<f4", offset=off).reshape(8, N[0] + 1, N[1] + 1, N[2] + 1)
    return dict(spacing=h[0], N=N, size=h[4:7], shift=h[7:10], delta=h[10:13], ucs=h[13:16],
                ncells=h[16:19], ewald_precision=ep, data=data)


@pytest.mark.parametrize("coulomb", [False, True])
def test_grid_file_round_trip(W, O, R, tmp_path, coulomb):
    """WriteVDWGrid / WriteCoulombGrid layout (header fields in order, then
    channel-major m, i, j, k floats), and reading the file back gives the
    same grid bit for bit."""
    fw = W.c1_framework()
    ctx = fw.context(charge_method=R.CHARGE_EWALD)
    if coulomb:
        g = ctx.MakeCoulombGrid(fw.unit_box, 0.7)
    else:
        g = ctx.MakeGrid(ctx.add_pseudo_atom(148.0, 3.73, 0.0), fw.unit_box, 0.7)
    path = str(tmp_path / "fw.grid")
    g.write(path)
    f = _parse(path, coulomb)
    data = g.download()
    assert f["spacing"] == 0.7
    assert tuple(f["N"]) == tuple(g.n)
    assert np.array_equal(f["size"], g.size) and np.array_equal(f["shift"], g.shift)
    assert np.array_equal(f["delta"], g.delta)
    assert np.array_equal(f["ucs"], [25.8, 25.8, 25.8])
    assert tuple(f["ncells"]) == (2, 2, 2)
    if coulomb:
        assert f["ewald_precision"] == 1e-6
    assert np.array_equal(f["data"], np.moveaxis(data, 3, 0))
    h = ctx.ReadGrid(path, fw.unit_box, coulomb=coulomb, n_unit_cells=(2, 2, 2))
    assert np.array_equal(h.download().view(np.uint32), data.view(np.uint32))
    assert h.coulomb == coulomb and h.spacing == 0.7
    pts = random_positions(fw.box, 500, 53)
    q = -0.35 if coulomb else None
    assert same_bits(h.interpolate(pts, q), g.interpolate(pts, q))
    # mismatches the reference rejects, and a damaged file
    with pytest.raises(R.RhgError) as e:
        ctx.ReadGrid(path, W.cell_box(25.9, 25.8, 25.8), coulomb=coulomb)
    assert e.value.code == R.ERR_INVALID
    if coulomb:
        with pytest.raises(R.RhgError) as e:
            ctx.ReadGrid(path, fw.unit_box, coulomb=True, n_unit_cells=(1, 2, 2))
        assert e.value.code == R.ERR_INVALID
    bad = str(tmp_path / "short.grid")
    open(bad, "wb").write(open(path, "rb").read()[:-4])
    with pytest.raises(R.RhgError) as e:
        ctx.ReadGrid(bad, fw.unit_box, coulomb=coulomb)
    assert e.value.code == R.ERR_IO
    with pytest.raises(R.RhgError) as e:
        ctx.ReadGrid(path, fw.unit_box, coulomb=not coulomb)
    assert e.value.code == R.ERR_IO


def test_grid_file_from_oracle(W, O, R, tmp_path):
    """A VDW grid file laid out as WriteVDWGrid writes it, from the oracle's
    node values, reads onto the device and interpolates as the oracle does."""
    import struct
    fw = W.c2_framework()
    sys_o = oracle_system(fw)
    ctx = fw.context()
    to = sys_o.add_pseudo_atom(148.0, 3.73, 0.0)
    spacing = 2.0
    geom = O.grid_geometry(fw.unit_box, spacing)
    N = geom[:3].astype(int)
    nodes = np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 8), dtype=np.float32)
    for i in range(N[0] + 1):
        for j in range(N[1] + 1):
            for k in range(N[2] + 1):
                nodes[i, j, k] = sys_o.grid_node(to, geom, i, j, k)
    inv, props = O.invert3x3(fw.unit_box)
    ucs = [np.linalg.norm(fw.unit_box[0:3]), np.linalg.norm(fw.unit_box[3:6]),
           np.linalg.norm(fw.unit_box[6:9])]
    hdr = struct.pack(_HDR, spacing, *N, *geom[3:6], *geom[6:9], *geom[9:12], *ucs, 1, 1, 1)
    path = str(tmp_path / "c2_CH4_shifted.grid")
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(np.moveaxis(nodes, 3, 0)).astype("<f4").tobytes())
    g = ctx.ReadGrid(path, fw.unit_box)
    assert np.array_equal(g.download().view(np.uint32), nodes.view(np.uint32))
    pts = random_positions(fw.box, 300, 59)
    u = g.interpolate(pts)
    for p, ug in zip(pts, u):
        ref = O.interpolate_grid(nodes, geom, O.BC_TRICLINIC, fw.unit_box, inv, p)
        assert abs(ug - ref) <= 1e-9 * max(1.0, abs(ref))


def test_c2_full_grid_against_oracle(W, O, R):
    """The C2 energy grid at its full size: methane probe over the triclinic
    26 A / 80 deg cell (1200 atoms) at 0.1 A spacing, 352 x 294 x 254 nodes
    (MakeGrid's node rule, src/grids.c:431-583).  At 10^4 seeded off-grid
    points the device interpolation of the FP64 build (FP32 node storage, as
    RASPA2 stores grids) equals the oracle's InterpolateVDWGrid over corner
    nodes computed on the fly within 1e-9 relative (max(1, |U|)); the mixed
    build keeps every capped node and agrees within 1e-5 relative where
    |U| > 1 K, 1e-3 K below."""
    fw = W.c2_framework()
    sys_o = oracle_system(fw)
    to = sys_o.add_pseudo_atom(148.0, 3.73, 0.0)
    geom = O.grid_geometry(fw.unit_box, 0.1)
    assert tuple(int(v) + 1 for v in geom[:3]) == (352, 294, 254)
    inv, _ = O.invert3x3(fw.unit_box)
    rng = np.random.Generator(np.random.PCG64(1201))
    pts = rng.uniform(0.0, 1.0, (10000, 3)) @ np.asarray(fw.unit_box).reshape(3, 3)
    ref = np.array([sys_o.interpolate_direct(to, geom, O.BC_TRICLINIC, fw.unit_box, inv, p) for p in pts])
    e2k = O.units()["energy_to_kelvin"]
    for prec in (R.PREC_FP64, R.PREC_MIXED):
        print(f"C2 0.1 A grid, precision {prec}", flush=True)
        ctx = fw.context(precision=prec)
        tg = ctx.add_pseudo_atom(148.0, 3.73, 0.0)
        g = ctx.MakeGrid(tg, fw.unit_box, 0.1)
        assert np.array_equal(g.geom, geom)
        u = g.interpolate(pts)
        capped = ref >= 1e7
        assert np.array_equal(u >= 1e7, capped)
        ok = ~capped
        assert ok.sum() > 1500  # the 1200-atom cell is dense: most points overlap
        if prec == R.PREC_FP64:
            assert np.max(np.abs(u[ok] - ref[ok]) / np.maximum(np.abs(ref[ok]), 1.0)) <= 1e-9
        else:
            err = np.abs(u[ok] - ref[ok]) * e2k
            big = np.abs(ref[ok]) * e2k > 1.0
            assert np.all(err[big] <= 1e-5 * np.abs(ref[ok][big]) * e2k)
            assert np.all(err[~big] <= 1e-3)
        g.close()
        ctx.close()
