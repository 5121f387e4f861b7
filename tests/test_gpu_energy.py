"""Energies at positions on the GPU against the oracle
(src/framework_energy.c:60-311)."""
import numpy as np
import pytest

from helpers import oracle_system, random_positions, same_bits

pytestmark = pytest.mark.gpu


def _pair(W, O, R, fw, charge_method, precision, probe_sites, cull=True, cell=2.5, images=True):
    sys_o = oracle_system(fw, charge_method=charge_method)
    ctx = fw.context(charge_method=charge_method, precision=precision)
    ctx.set_culling(cull, cell, images)
    to = [sys_o.add_pseudo_atom(*s) for s in probe_sites]
    tg = [ctx.add_pseudo_atom(*s) for s in probe_sites]
    return sys_o, ctx, to, tg


FRAMEWORKS = {"c0": lambda W: W.c0_framework(), "c1": lambda W: W.c1_framework(),
              "c2": lambda W: W.c2_framework()}


@pytest.mark.parametrize("cull,cell", [(False, 2.5), (True, 2.5), (True, 4.0)])
@pytest.mark.parametrize("which", ["c0", "c1", "c2"])
def test_vdw_fp64_bit_exact(W, O, R, which, cull, cell):
    """Plain FP64 LJ energies equal the sequential reference loop bit for bit
    (rectangular C0/C1 cells, triclinic C2 cell), with and without cell-list
    culling."""
    fw = FRAMEWORKS[which](W)
    sys_o, ctx, to, tg = _pair(W, O, R, fw, O.CHARGE_NONE, R.PREC_FP64, [(148.0, 3.73, 0.0)],
                               cull, cell)
    xyz = random_positions(fw.box, 3000, 7)
    uv, uc, ov = ctx.energy_at_positions(xyz, tg[0])
    uv_o, _ = sys_o.energy_at_many(to[0], xyz)
    assert same_bits(uv, uv_o)
    assert np.array_equal(ov, uv_o >= 1e7)
    assert np.all(uc == 0.0)


@pytest.mark.parametrize("cull,images", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("which", ["c0", "c1", "c2"])
def test_vdw_mixed_tolerance_and_flags(W, O, R, which, cull, images):
    """Mixed precision: overlap flags identical, relative error <= 1e-5 on
    non-overlapping sites (|U| > 1 K), absolute <= 1e-3 K otherwise; per-pair
    minimum image (all pairs or cell lists) and image lists."""
    fw = FRAMEWORKS[which](W)
    sys_o, ctx, to, tg = _pair(W, O, R, fw, O.CHARGE_NONE, R.PREC_MIXED, [(148.0, 3.73, 0.0)],
                               cull, images=images)
    xyz = random_positions(fw.box, 3000, 8)
    uv, _, ov = ctx.energy_at_positions(xyz, tg[0])
    uv_o, _ = sys_o.energy_at_many(to[0], xyz)
    assert np.array_equal(ov, uv_o >= 1e7)
    ok = uv_o < 1e7
    e2k = O.units()["energy_to_kelvin"]
    err = np.abs(uv[ok] - uv_o[ok]) * e2k
    big = np.abs(uv_o[ok]) * e2k > 1.0
    assert np.all(err[big] / (np.abs(uv_o[ok][big]) * e2k) <= 1e-5)
    assert np.all(err[~big] <= 1e-3)


@pytest.mark.parametrize("cull,images", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("precision", [0, 1])
def test_charge_energy(W, O, R, precision, cull, images):
    """Real-space Ewald energy of a charged bead in the 2x2x2 C1 supercell."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg = _pair(W, O, R, fw, O.CHARGE_EWALD, precision, [(27.0, 2.80, 0.70)],
                               cull, images=images)
    xyz = random_positions(fw.box, 1500, 9)
    uv, uc, ov = ctx.energy_at_positions(xyz, tg[0])
    uv_o, uc_o = sys_o.energy_at_many(to[0], xyz)
    ok = (uv_o < 1e7) & (uc_o < 1e7)
    assert np.array_equal(ov, ~ok)
    err = np.abs(uc[ok] - uc_o[ok])
    if precision == 0:
        assert np.max(err / np.maximum(np.abs(uc_o[ok]), 1.0)) <= 1e-11
        assert same_bits(uv, uv_o)
    else:
        # FP32 pair terms: the error of a sum is bounded by a few FP32 ulps of
        # the summed term magnitudes (cancellation between charges of both
        # signs makes a bound relative to the sum itself meaningless)
        scale = _coulomb_abs_sum(fw, xyz[ok], 0.70, ctx.info()["alpha"],
                                 O.units()["coulomb"])
        assert np.max(err / scale) <= 1e-6


def _coulomb_abs_sum(fw, xyz, qa, alpha, C, rc=12.0):
    """sum_j |C qa q_j erfc(alpha r)/r| over r < rc (cubic cell, minimum image)."""
    from scipy.special import erfc
    L = float(np.asarray(fw.box).reshape(9)[0])
    out = np.empty(len(xyz))
    for i, p in enumerate(xyz):
        d = fw.xyz - p
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        m = r < rc
        out[i] = np.sum(np.abs(C * qa * fw.charge[m] * erfc(alpha * r[m]) / r[m]))
    return out


def test_scalar_shims(W, O, R):
    """The legacy call shapes (one position per call)."""
    fw = W.c1_framework()
    sys_o, ctx, to, tg = _pair(W, O, R, fw, O.CHARGE_EWALD, R.PREC_FP64, [(27.0, 2.80, 0.70)])
    p = random_positions(fw.box, 3, 10)
    for pos in p:
        u = ctx.CalculateFrameworkVDWEnergyAtPosition(pos, tg[0])
        assert u == sys_o.vdw_at(to[0], pos)
        ucc, ucbd = ctx.CalculateFrameworkChargeEnergyAtPosition(pos, tg[0])
        assert ucbd == 0.0
        assert abs(ucc - sys_o.charge_at(to[0], pos)) <= 1e-11 * max(1.0, abs(ucc))


def _clustered_positions(box, n, seed):
    """Half of the points in a 1.5 A cube, a quarter in a 6 A cube, the rest
    uniform: cells holding one to several hundred sites, so the mixed
    image-list kernels run tiles of every size (one to four sites per lane,
    several tiles per cell)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.asarray(box).reshape(3, 3)
    c1, c2 = np.array([0.31, 0.42, 0.53]) @ m, np.array([0.71, 0.22, 0.64]) @ m
    a = c1 + rng.uniform(-0.75, 0.75, (n // 2, 3))
    b = c2 + rng.uniform(-3.0, 3.0, (n // 4, 3))
    u = rng.uniform(0.0, 1.0, (n - n // 2 - n // 4, 3)) @ m
    return np.concatenate([a, b, u])


@pytest.mark.parametrize("charged", [False, True])
def test_mixed_dense_cells(W, O, R, charged):
    """Mixed precision with hundreds of sites per cell (C1 supercell, LJ and
    LJ + real-space Ewald): overlap flags identical to the oracle's, LJ within
    1e-5 relative (|U| > 1 K) or 1e-3 K, Coulomb within 1e-6 of the summed
    term magnitudes."""
    fw = W.c1_framework()
    site = (27.0, 2.80, 0.70) if charged else (148.0, 3.73, 0.0)
    cm = O.CHARGE_EWALD if charged else O.CHARGE_NONE
    sys_o, ctx, to, tg = _pair(W, O, R, fw, cm, R.PREC_MIXED, [site], True, 2.5, True)
    xyz = _clustered_positions(fw.box, 6000, 21)
    uv, uc, ov = ctx.energy_at_positions(xyz, tg[0])
    uv_o, uc_o = sys_o.energy_at_many(to[0], xyz)
    bad = (uv_o >= 1e7) | ((uc_o >= 1e7) if charged else False)
    assert np.array_equal(ov, bad)
    ok = ~bad
    assert ok.sum() > 500
    e2k = O.units()["energy_to_kelvin"]
    err = np.abs(uv[ok] - uv_o[ok]) * e2k
    big = np.abs(uv_o[ok]) * e2k > 1.0
    assert np.all(err[big] / (np.abs(uv_o[ok][big]) * e2k) <= 1e-5)
    assert np.all(err[~big] <= 1e-3)
    if charged:
        scale = _coulomb_abs_sum(fw, xyz[ok], 0.70, ctx.info()["alpha"], O.units()["coulomb"])
        assert np.max(np.abs(uc[ok] - uc_o[ok]) / scale) <= 1e-6


def _threshold_positions(sys_o, t, fw, n_dir, seed):
    """Positions whose FP64 VDW energy lies within a few 1e-6 (relative) of
    the overlap threshold 1e7, on both sides: bisection on the oracle along
    rays from framework atoms (the energy falls steeply with the distance
    there)."""
    rng = np.random.default_rng(seed)
    atoms = fw.xyz[rng.choice(fw.n, n_dir, replace=False)]
    d = rng.normal(size=(n_dir, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]

    def excess(r):
        uv, _ = sys_o.energy_at_many(t, atoms + r[:, None] * d)
        return uv - 1e7

    lo, hi = np.full(n_dir, 0.3), np.full(n_dir, 3.0)
    keep = (excess(lo) > 0) & (excess(hi) < 0)
    lo, hi, atoms, d = lo[keep], hi[keep], atoms[keep], d[keep]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        above = excess(mid) > 0
        lo, hi = np.where(above, mid, lo), np.where(above, hi, mid)
    r0 = 0.5 * (lo + hi)
    steps = np.array([-3e-7, -1e-7, -3e-8, 0.0, 3e-8, 1e-7, 3e-7])
    r = (r0[:, None] * (1.0 + steps[None, :])).ravel()
    return np.repeat(atoms, len(steps), axis=0) + r[:, None] * np.repeat(d, len(steps), axis=0)


@pytest.mark.parametrize("charged", [False, True])
def test_guard_band_sites_recomputed_exactly(W, O, R, charged):
    """Mixed precision at positions whose energy straddles the overlap
    threshold within 1e-5: the image-list kernel lists them and the exact
    FP64 recompute gives flags equal to the oracle's and energies bitwise
    equal to the FP64 path's (C1 supercell, image lists)."""
    fw = W.c1_framework()
    cm = O.CHARGE_EWALD if charged else O.CHARGE_NONE
    probe = [(27.0, 2.80, 0.70)] if charged else [(148.0, 3.73, 0.0)]
    sys_o, ctx, to, tg = _pair(W, O, R, fw, cm, R.PREC_MIXED, probe, True, 2.5, True)
    _, ref, _, rg = _pair(W, O, R, fw, cm, R.PREC_FP64, probe, True, 2.5, True)
    xyz = _threshold_positions(sys_o, to[0], fw, 96, 21)
    xyz = np.concatenate([xyz, random_positions(fw.box, 2000, 22)])
    print(f"guard band: {len(xyz)} positions, charged={charged}", flush=True)
    uv, uc, ov = ctx.energy_at_positions(xyz, tg[0])
    uv64, uc64, ov64 = ref.energy_at_positions(xyz, rg[0])
    uv_o, uc_o = sys_o.energy_at_many(to[0], xyz)
    flag_o = (uv_o >= 1e7) | ((uc_o >= 1e7) if charged else False)
    assert np.array_equal(ov, flag_o) and np.array_equal(ov64, flag_o)
    band = np.abs(uv_o / 1e7 - 1.0) < 1e-5
    assert band.sum() >= 100 and 0 < flag_o[band].sum() < band.sum()
    assert same_bits(uv[band], uv64[band])
    if charged:
        assert same_bits(uc[band], uc64[band])
    ctx.close()
    ref.close()


def test_coulomb_only_lists_match_lj_coulomb_lists(W, O, R):
    """Coulomb-only image lists (two 32-site tiles per wave, one cell per
    half-wave) sum each site's real-space Ewald terms exactly as the LJ +
    Coulomb kernel does: with the LJ energy served from a grid, the positions'
    Coulomb energies equal those of the direct LJ + Coulomb pass bit for bit
    (scattered and crowded cells; sites near the overlap threshold, which the
    direct pass recomputes in FP64 for their LJ sum, left out)."""
    fw = W.c1_framework()
    site = (27.0, 2.80, 0.70)
    _, ctx_a, _, ta = _pair(W, O, R, fw, O.CHARGE_EWALD, R.PREC_MIXED, [site], True, 2.5, True)
    _, ctx_b, _, tb = _pair(W, O, R, fw, O.CHARGE_EWALD, R.PREC_MIXED, [site], True, 2.5, True)
    g = ctx_b.MakeGrid(tb[0], fw.unit_box, 0.5)
    ctx_b.use_grid(tb[0], g)
    xyz = np.concatenate([random_positions(fw.box, 3000, 31), _clustered_positions(fw.box, 6000, 32)])
    uv_a, uc_a, _ = ctx_a.energy_at_positions(xyz, ta[0])
    _, uc_b, _ = ctx_b.energy_at_positions(xyz, tb[0])
    keep = (uv_a < 1e6) & (uc_a < 1e6)
    assert keep.sum() > 2000
    assert same_bits(uc_a[keep], uc_b[keep])
