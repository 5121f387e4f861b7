"""Pins of the oracle by what the mathematics fixes: closed forms, invariants,
brute force on small inputs, textbook constants.  Each test is chosen so that
a plausible slip (dropped term, wrong sign or index, transposed operand)
fails it."""
import math

import numpy as np
import pytest
from scipy.special import erf, erfc

from helpers import oracle_system, random_positions


def _one_atom_system(O, box, bc, eps=50.0, sig=3.0, q=0.0, cm=0, rc=12.0):
    return O.System(box, bc, [0.0], [0.0], [0.0], [q], [eps], [sig], cutoff_vdw=rc,
                    shifted=True, charge_method=cm, cutoff_charge=rc)


def test_lj_closed_form(O):
    """U(r) = 4 eps[(s/r)^12-(s/r)^6] - U(rc) with LB-mixed eps, sigma;
    zero beyond the cutoff; minimum -eps - U(rc) at 2^(1/6) sigma."""
    box, inv, props = O.cell_make(30.0, 30.0, 30.0)
    s = _one_atom_system(O, box, O.BC_RECTANGULAR)
    t = s.add_pseudo_atom(148.0, 3.73)
    k2e = O.units()["kelvin_to_energy"]
    eps = math.sqrt(50.0 * k2e * 148.0 * k2e)
    sig = 0.5 * (3.0 + 3.73)

    def lj(r):
        return 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)

    shift = lj(12.0)
    for r in (2.5, 3.0, 3.7, 5.0, 8.0, 11.99):
        u = s.vdw_at(t, [r, 0.0, 0.0])
        assert abs(u - (lj(r) - shift)) <= 1e-12 * max(1.0, abs(lj(r)))
    assert s.vdw_at(t, [12.0001, 0, 0]) == 0.0
    rmin = 2 ** (1 / 6) * sig
    assert abs(s.vdw_at(t, [0, rmin, 0]) - (-eps - shift)) < 1e-12
    assert abs(s.vdw_at(t, [0, 0, 11.999999])) < 1e-8 * eps


@pytest.mark.parametrize("bc", [1, 2])
def test_lattice_translation_invariance(O, W, bc):
    """Energies do not change when the probe moves by any lattice vector
    (minimum image in rectangular and triclinic form)."""
    fw = W.c2_framework() if bc == 2 else W.c0_framework()
    s = oracle_system(fw)
    t = s.add_pseudo_atom(148.0, 3.73)
    m = np.asarray(fw.box).reshape(3, 3)
    p = random_positions(fw.box, 20, 3)
    for shift in ([1, 0, 0], [0, -1, 0], [0, 0, 2], [1, 1, -1]):
        d = np.array(shift) @ m
        for x in p:
            u0, u1 = s.vdw_at(t, x), s.vdw_at(t, x + d)
            assert abs(u0 - u1) <= 1e-9 * max(1.0, abs(u0))


def test_triclinic_equals_rectangular_on_orthorhombic(O, W):
    fw = W.c0_framework()
    s1 = oracle_system(fw)
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    s2 = O.System(box, O.BC_TRICLINIC, x, y, z, q, eps, sig, mass=mass)
    t1, t2 = s1.add_pseudo_atom(148.0, 3.73), s2.add_pseudo_atom(148.0, 3.73)
    for p in random_positions(fw.box, 50, 4):
        u1, u2 = s1.vdw_at(t1, p), s2.vdw_at(t2, p)
        assert abs(u1 - u2) <= 1e-9 * max(1.0, abs(u1))


def test_minimum_image_brute_force(O):
    """All periodic images within the cutoff, enumerated explicitly, equal
    the minimum-image sum when the cutoff is below half the perpendicular
    width (triclinic cell)."""
    box, inv, props = O.cell_make(25.0, 26.0, 27.0, 75.0, 100.0, 85.0)
    rng = np.random.default_rng(2)
    m = box.reshape(3, 3)
    n = 40
    xyz = rng.uniform(0, 1, (n, 3)) @ m
    eps, sig = rng.uniform(20, 80, n), rng.uniform(2.5, 3.5, n)
    rc = 11.0
    assert props[3:6].min() > 2 * rc
    s = O.System(box, O.BC_TRICLINIC, xyz[:, 0], xyz[:, 1], xyz[:, 2], np.zeros(n), eps, sig,
                 cutoff_vdw=rc)
    t = s.add_pseudo_atom(148.0, 3.73)
    k2e = O.units()["kelvin_to_energy"]
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]) @ m
    for p in rng.uniform(0, 1, (10, 3)) @ m:
        tot = 0.0
        for a in range(n):
            e = math.sqrt(eps[a] * k2e * 148.0 * k2e)
            sg = 0.5 * (sig[a] + 3.73)
            sh = 4 * e * ((sg / rc) ** 12 - (sg / rc) ** 6)
            for d in shifts:
                r = np.linalg.norm(p - (xyz[a] + d))
                if r < rc:
                    tot += 4 * e * ((sg / r) ** 12 - (sg / r) ** 6) - sh
        u = s.vdw_at(t, p)
        assert abs(u - tot) <= 1e-9 * max(1.0, abs(tot))


def _nacl(O, reps=5, a=5.64, precision=1e-6, rc=12.0):
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    na = base
    cl = base + np.array([.5, 0, 0])
    frac, q = [], []
    for i in range(reps):
        for j in range(reps):
            for k in range(reps):
                off = np.array([i, j, k])
                frac += list(na + off) + list(cl + off)
                q += [1.0] * 4 + [-1.0] * 4
    L = reps * a
    box, inv, props = O.cell_make(L, L, L)
    xyz = np.array(frac) * a
    n = len(q)
    s = O.System(box, O.BC_RECTANGULAR, xyz[:, 0], xyz[:, 1], xyz[:, 2], q, np.zeros(n),
                 np.full(n, 3.0), charge_method=O.CHARGE_EWALD, cutoff_charge=rc,
                 cutoff_vdw=rc, ewald_precision=precision)
    return s, n, a


@pytest.mark.parametrize("precision,tol", [(1e-6, 2e-5), (1e-9, 5e-8)])
def test_ewald_madelung_nacl(O, precision, tol):
    """Real + Fourier - self Ewald energy of rock salt equals the Madelung
    energy -M N C/(2 r0), M = 1.747564594633, to the requested precision."""
    s, n, a = _nacl(O, precision=precision)
    C = O.units()["coulomb"]
    e = s.real_energy_framework() + s.fourier_self_framework() - s.self_energy_framework()
    ref = -1.747564594633 * n * C / (2 * (a / 2))
    assert abs(e / ref - 1.0) < tol


def test_ewald_alpha_independence(O):
    """The total Ewald energy does not depend on the splitting parameter."""
    s1, n, a = _nacl(O, precision=1e-7)
    s2, _, _ = _nacl(O, precision=1e-9)
    assert s1.info()["alpha"] != s2.info()["alpha"]
    e1 = s1.real_energy_framework() + s1.fourier_self_framework() - s1.self_energy_framework()
    e2 = s2.real_energy_framework() + s2.fourier_self_framework() - s2.self_energy_framework()
    assert abs(e1 / e2 - 1.0) < 2e-6


def test_fourier_delta_is_energy_difference(O, W):
    """CalculateEwaldFourierAdsorbate's host + adsorbate deltas plus the
    excluded term equal the change of sum_k kf |S(k)|^2 when the molecule's
    charges join the framework, minus the molecule's self energy."""
    fw = W.c1_framework(n_rep=1)
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    types = [s.add_pseudo_atom(*site) for site in W.CO2.sites]
    c = np.array([5.0, 6.0, 7.0])
    v = np.array([0.3, -0.5, 0.81])
    v /= np.linalg.norm(v)
    mol = np.array([c + 1.16 * v, c, c - 1.16 * v])
    out = s.fourier_adsorbate(types, mol)
    qm = [site[2] for site in W.CO2.sites]
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    s2 = O.System(box, bc, np.r_[x, mol[:, 0]], np.r_[y, mol[:, 1]], np.r_[z, mol[:, 2]],
                  np.r_[q, qm], np.r_[eps, [1, 1, 1]], np.r_[sig, [1, 1, 1]],
                  charge_method=O.CHARGE_EWALD, ewald_precision=1e-6)
    assert s2.info()["nk"] == s.info()["nk"]
    d_recip = s2.fourier_self_framework() - s.fourier_self_framework()
    C, alpha = O.units()["coulomb"], s.info()["alpha"]
    self_mol = sum(C * qq * qq * alpha / math.sqrt(math.pi) for qq in qm)
    assert abs(out[2] - self_mol) < 1e-9 * self_mol
    excl = 0.0
    for i in range(3):
        for j in range(i + 1, 3):
            r = np.linalg.norm(mol[i] - mol[j])
            excl += C * qm[i] * qm[j] * erf(alpha * r) / r
    assert abs(out[3] - excl) < 1e-9 * abs(excl)
    assert abs((out[0] + out[1] + out[3]) - (d_recip - self_mol)) < 1e-8 * abs(self_mol)


def test_charge_energy_closed_form(O):
    """Real-space Ewald pair term C qa qb erfc(alpha r)/r (scipy erfc)."""
    box, inv, props = O.cell_make(30.0, 30.0, 30.0)
    s = _one_atom_system(O, box, O.BC_RECTANGULAR, q=-0.4, cm=O.CHARGE_EWALD)
    t = s.add_pseudo_atom(27.0, 2.8, 0.7)
    C, alpha = O.units()["coulomb"], s.info()["alpha"]
    for r in (1.0, 2.5, 7.0, 11.9):
        u = s.charge_at(t, [0, r, 0])
        ref = C * 0.7 * -0.4 * erfc(alpha * r) / r
        assert abs(u - ref) <= 1e-13 * abs(ref)
    assert s.charge_at(t, [0, 12.01, 0]) == 0.0


def _tricubic_grid(seed):
    """A 2x2x2-cell grid over [0,2]^3 (unit spacing) holding a random
    tricubic polynomial and its channel derivatives, and the polynomial."""
    rng = np.random.default_rng(seed)
    coef = rng.normal(size=(4, 4, 4))

    def f_and_derivs(x, y, z):
        out = np.zeros(8)
        p = [np.array([x ** i for i in range(4)]), np.array([y ** i for i in range(4)]),
             np.array([z ** i for i in range(4)])]
        dp = [np.array([i * x ** (i - 1) if i else 0.0 for i in range(4)]),
              np.array([i * y ** (i - 1) if i else 0.0 for i in range(4)]),
              np.array([i * z ** (i - 1) if i else 0.0 for i in range(4)])]
        orders = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1),
                  (1, 1, 1)]
        for ch, (a, b, c) in enumerate(orders):
            px = dp[0] if a else p[0]
            py = dp[1] if b else p[1]
            pz = dp[2] if c else p[2]
            out[ch] = np.einsum("ijk,i,j,k->", coef, px, py, pz)
        return out

    geom = np.array([2, 2, 2, 2.0, 2.0, 2.0, 0, 0, 0, 1.0, 1.0, 1.0])
    grid = np.zeros((3, 3, 3, 8), dtype=np.float32)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                grid[i, j, k] = f_and_derivs(i, j, k)
    box = np.array([2.0, 0, 0, 0, 2.0, 0, 0, 0, 2.0])
    inv = np.array([0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5])
    return grid, geom, box, inv, f_and_derivs


def test_lekien_marsden_matrix(O):
    """The generated 64x64 matrix is integer, inverts the constraint matrix,
    and interpolates any tricubic polynomial exactly."""
    C = O.lm_coeff()
    assert C.dtype == np.int32 and np.abs(C).max() <= 100
    grid, geom, box, inv, f_and_derivs = _tricubic_grid(9)
    rng = np.random.default_rng(9)
    g64 = grid.astype(np.float64)
    for p in rng.uniform(0.05, 1.95, (50, 3)):
        val = O.interpolate_grid(grid, geom, O.BC_RECTANGULAR, box, inv, p)
        # exact tricubic of the float32-rounded node data: compare with the
        # polynomial within the float32 rounding of the inputs
        ref = f_and_derivs(*p)[0]
        assert abs(val - ref) <= 1e-5 * max(1.0, np.abs(g64).max())


def test_coulomb_interpolation(O):
    """InterpolateCoulombGrid: the same tricubic, times the charge, and no
    overlap short-cut (a node above 1e10 still interpolates)."""
    grid, geom, box, inv, f_and_derivs = _tricubic_grid(10)
    rng = np.random.default_rng(10)
    g64 = grid.astype(np.float64)
    for p in rng.uniform(0.05, 1.95, (30, 3)):
        ref = f_and_derivs(*p)[0]
        v1 = O.interpolate_coulomb_grid(grid, geom, O.BC_RECTANGULAR, box, inv, p, 1.0)
        assert abs(v1 - ref) <= 1e-5 * max(1.0, np.abs(g64).max())
        v2 = O.interpolate_coulomb_grid(grid, geom, O.BC_RECTANGULAR, box, inv, p, -0.35)
        assert v2 == -0.35 * v1
    big = grid.copy()
    big[1, 1, 1, 0] = 2e10
    p = np.array([0.5, 0.5, 0.5])
    assert O.interpolate_grid(big, geom, O.BC_RECTANGULAR, box, inv, p) == 1e8
    v = O.interpolate_coulomb_grid(big, geom, O.BC_RECTANGULAR, box, inv, p, 1.0)
    # the corner (1,1,1) enters with the weight h(0.5)^3 = 1/8 of its value
    v0 = O.interpolate_coulomb_grid(grid, geom, O.BC_RECTANGULAR, box, inv, p, 1.0)
    assert abs((v - v0) - (2e10 - float(grid[1, 1, 1, 0])) / 8.0) <= 1e-6 * 2e10 / 8.0


def test_real_derivatives_finite_differences(O, W):
    """CalculateDerivativesAtPositionReal for a unit charge: the value is the
    real-space energy of a unit-charge probe, the gradient, the mixed second
    derivatives and the third derivative match central differences, the
    Hessian's diagonal too, and the return value is the nearest in-range
    atom."""
    fw = W.c1_framework(n_rep=1)
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    t = s.add_pseudo_atom(0.0, 1.0, 1.0)
    h = 1e-3
    xyz = np.stack([fw.xyz[:, 0], fw.xyz[:, 1], fw.xyz[:, 2]], axis=1)
    L = fw.box[0]
    checked = 0
    for p in random_positions(fw.box, 12, 23):
        d, rmin = s.derivatives_real_at(p)
        dr = p - xyz
        dr -= L * np.round(dr / L)
        r = np.sqrt((dr ** 2).sum(axis=1))
        assert rmin == pytest.approx(r[r < 12.0].min(), rel=1e-14)
        if rmin < 1.2:
            continue
        checked += 1
        u = s.charge_at(t, p)
        terms = O.units()["coulomb"] * np.abs(fw.charge[r < 12.0]) / r[r < 12.0]
        assert abs(d[0] - u) <= 1e-13 * terms.sum()
        U = lambda q: s.derivatives_real_at(q)[0][0]
        ex, ey, ez = np.eye(3) * h
        g = [(U(p + e) - U(p - e)) / (2 * h) for e in (ex, ey, ez)]
        scale = max(1.0, np.abs(d[1:4]).max())
        assert np.allclose(d[1:4], g, atol=1e-5 * scale)
        for (a, b, col) in ((ex, ey, 5), (ex, ez, 6), (ey, ez, 9)):
            hab = (U(p + a + b) - U(p + a - b) - U(p - a + b) + U(p - a - b)) / (4 * h * h)
            assert abs(d[col] - hab) <= 1e-3 * max(1.0, abs(hab))
        hxx = (U(p + ex) - 2 * U(p) + U(p - ex)) / (h * h)
        assert abs(d[4] - hxx) <= 1e-3 * max(1.0, abs(hxx))
        gxy = lambda q: s.derivatives_real_at(q)[0][5]
        t3 = (gxy(p + ez) - gxy(p - ez)) / (2 * h)
        assert abs(d[13] - t3) <= 1e-3 * max(1.0, abs(t3))
    assert checked >= 6


def test_coulomb_grid_node_cap(O):
    """The Coulomb grid's cap: a node closer than 1 A to an atom is capped
    although its energy is small, a node whose energy exceeds 1e7 is capped
    although no atom is that close, and an uncapped node stores the
    derivatives scaled by the node spacing."""
    box, inv, props = O.cell_make(30.0, 30.0, 30.0)
    for q, r, capped in ((0.5, 0.8, True), (0.5, 1.5, False), (200.0, 1.1, True)):
        s = O.System(box, O.BC_RECTANGULAR, [0.0], [0.0], [0.0], [q], [50.0], [3.0],
                     cutoff_vdw=12.0, shifted=True, charge_method=O.CHARGE_EWALD, cutoff_charge=12.0)
        # a 2-cell grid whose node (1, 0, 0) lies at distance r from the atom
        geom = np.array([2, 2, 2, 2 * r, 2 * r, 2 * r, 0, 0, 0, r, r, r])
        node = s.coulomb_grid_node(geom, 1, 0, 0)
        d, rmin = s.derivatives_real_at([r, 0.0, 0.0])
        assert rmin == r
        if capped:
            # exactly one of the two rules applies in each capped case
            assert (d[0] > 1e7) != (rmin < 1.0)
            assert node[0] == np.float32(2e7)
            assert np.all(node[4:] == 0.0)
        else:
            assert d[0] < 1e7 and node[0] == np.float32(d[0])
            assert node[1] == np.float32(np.float64(np.float32(d[1])) * r)
            assert node[4] == np.float32(np.float64(np.float32(d[5])) * (r * r))


def test_grid_derivatives_finite_differences(O, W):
    """CalculateDerivativesAtPositionVDW against central differences of the
    energy: gradient, the mixed second derivatives and the third derivative."""
    fw = W.c2_framework()
    s = oracle_system(fw)
    t = s.add_pseudo_atom(148.0, 3.73)
    h = 1e-3
    for p in random_positions(fw.box, 6, 17):
        d = s.derivatives_vdw_at(t, p)
        if d[0] > 1e5:
            continue
        U = lambda q: s.derivatives_vdw_at(t, q)[0]
        ex, ey, ez = np.eye(3) * h
        g = [(U(p + e) - U(p - e)) / (2 * h) for e in (ex, ey, ez)]
        scale = max(1.0, np.abs(d[1:4]).max())
        assert np.allclose(d[1:4], g, atol=1e-5 * scale)
        hxy = (U(p + ex + ey) - U(p + ex - ey) - U(p - ex + ey) + U(p - ex - ey)) / (4 * h * h)
        assert abs(d[5] - hxy) <= 1e-3 * max(1.0, abs(hxy))
        gxyz = lambda q: s.derivatives_vdw_at(t, q)[5]
        t3 = (gxyz(p + ez) - gxyz(p - ez)) / (2 * h)
        assert abs(d[13] - t3) <= 1e-3 * max(1.0, abs(t3))


def test_widom_without_framework_interaction(O, W):
    """With epsilon = 0 for every framework atom and no charges every trial
    energy is 0: W = 1 for every insertion and <U> = 0 (CBMC weights are
    normalised by the number of trials)."""
    fw = W.c0_framework()
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    s = O.System(box, bc, x, y, z, q, np.zeros_like(eps), sig, mass=mass)
    types = [s.add_pseudo_atom(e, sg, 0.0) for e, sg, _ in W.N2LIKE.sites]
    mol = O.molecule(types, W.N2LIKE.body, start=0, k_fb=7, k_or=5)
    rec = O.widom_run_stream(s, mol, 3, 0, 200)
    assert np.all(rec[:, 0] == 1.0)
    assert np.all(rec[:, 1] == 0.0)


def test_stream_trial_geometry(O, W):
    """Counter-stream trials: first beads inside the cell, rotated bodies
    keep their bond lengths and shape (orthonormal Shoemake rotation)."""
    fw = W.c1_framework(n_rep=1)
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    types = [s.add_pseudo_atom(*site) for site in W.CO2.sites]
    mol = O.molecule(types, W.CO2.body, start=1, k_fb=3, k_or=10)
    inv = np.linalg.inv(np.asarray(fw.box).reshape(3, 3))
    for i in range(50):
        fb, cord, usel = O.stream_trials(s, mol, 99, i)
        f = fb @ inv
        assert np.all((f >= 0) & (f <= 1))
        assert np.all((usel >= 0) & (usel <= 1))
        for o in range(10):
            assert np.allclose(cord[o, 1], 0.0, atol=0)
            assert abs(np.linalg.norm(cord[o, 0]) - 1.16) < 1e-14
            assert np.allclose(cord[o, 0], -cord[o, 2], atol=1e-15)


def test_stream_uniform_moments(O):
    u = np.concatenate([O.stream_uniforms(5, i, 100) for i in range(2000)])
    assert abs(u.mean() - 0.5) < 0.003
    assert abs(u.var() - 1 / 12) < 0.002
    assert len(np.unique(u)) == len(u)


def test_block_of_cycles(O):
    b = O.block_of_cycles(100)
    assert np.array_equal(np.bincount(b), [20] * 5)
    b = O.block_of_cycles(7)
    assert b[0] == 0 and b[-1] == 4 and np.all(np.diff(b) >= 0)


def test_mt_driver_matches_stream_statistically(O):
    """RASPA's global MT19937-64 draw order and the per-insertion counter
    stream sample the same distribution: He <W> in IRMOF-1 agrees."""
    import json, os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "irmof1.json")))
    a, b, c, al, be, ga = G["cell"]
    box, inv, props = O.cell_make(a, b, c, al, be, ga)
    frac = np.array([at["frac"] for at in G["atoms"]])
    xyz = frac @ box.reshape(3, 3)
    n = len(frac)
    s = O.System(box, O.BC_TRICLINIC, xyz[:, 0], xyz[:, 1], xyz[:, 2], np.zeros(n),
                 [at["eps_k"] for at in G["atoms"]], [at["sigma"] for at in G["atoms"]],
                 mass=[at["mass"] for at in G["atoms"]], temperature=298.0)
    he = s.add_pseudo_atom(10.9, 2.64)
    mol = O.molecule([he], [[0, 0, 0]], start=0, k_fb=10, k_or=10)
    n_ins, blocks = O.widom_run_mt(s, mol, G["output"]["seed"], 500, 20)
    assert n_ins == 10000
    w_mt = blocks[:, 1].sum() / blocks[:, 0].sum()
    rec = O.widom_run_stream(s, mol, 4, 0, 10000)
    w_st = rec[:, 0].mean()
    se = rec[:, 0].std() / np.sqrt(10000)
    assert abs(w_mt - w_st) < 5 * np.sqrt(2) * se


def test_mt_trials_record_replays(O, W):
    """orc_widom_run_mt_trials records everything an insertion of RASPA2's
    MT19937-64 driver drew: replaying the recorded first-bead positions,
    rotated offsets and selection uniforms through the host-trial path gives
    every record again, bit for bit, and the driver's block sums are the
    records' sums (helium in IRMOF-1 with 10 first-bead trials; CO2 with
    Ewald in the 25.8 A charged cell)."""
    import json, os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "irmof1.json")))
    a, b, c, al, be, ga = G["cell"]
    box, inv, props = O.cell_make(a, b, c, al, be, ga)
    frac = np.array([at["frac"] for at in G["atoms"]])
    xyz = frac @ box.reshape(3, 3)
    n = len(frac)
    s = O.System(box, O.BC_TRICLINIC, xyz[:, 0], xyz[:, 1], xyz[:, 2], np.zeros(n),
                 [at["eps_k"] for at in G["atoms"]], [at["sigma"] for at in G["atoms"]],
                 mass=[at["mass"] for at in G["atoms"]], temperature=298.0)
    he = s.add_pseudo_atom(10.9, 2.64)
    cases = [(s, O.molecule([he], [[0, 0, 0]], start=0, k_fb=10, k_or=10), 40)]
    fw = W.c1_framework(n_rep=1)
    s2 = O.System(*fw.oracle_args()[:8], mass=fw.mass, charge_method=O.CHARGE_EWALD)
    ty = [s2.add_pseudo_atom(*site) for site in W.CO2.sites]
    cases.append((s2, O.molecule(ty, W.CO2.body, start=W.CO2.start, k_fb=3, k_or=10), 20))
    for sys_o, mol, n_cycles in cases:
        n_ins, blocks, rec, fb, cord, usel = O.widom_run_mt_trials(sys_o, mol, 1417341545, n_cycles, 20)
        assert n_ins == 20 * n_cycles == len(rec)
        again = np.array([O.widom_insertion_trials(sys_o, mol, fb[i], cord[i] if mol.n_sites > 1 else
                                                   np.zeros((mol.k_or, 1, 3)), usel[i]) for i in range(n_ins)])
        assert np.array_equal(again, rec)
        assert (rec[:, 13] == 4).sum() > 10
        ok = rec[:, 6] == 0
        assert abs(rec[ok, 0].sum() - blocks[:, 1].sum()) <= 1e-12 * blocks[:, 1].sum()
        assert blocks[:, 0].sum() == n_ins


def test_sincos_portable_against_libm(O):
    """The portable sin/cos of the Shoemake angles (written identically for
    the device) against glibc's sin and cos on [0, 2 pi]: absolute error at
    most 2^-53 (half an ulp of 1), and at most one ulp of the value wherever
    |value| > 1e-3; the reduction is exact at multiples of pi/2."""
    import ctypes
    f = O.lib().orc_sincos_portable
    f.restype = None
    f.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0.0, 2.0 * np.pi, 50000), np.arange(5) * (np.pi / 2),
                        np.nextafter(np.arange(1, 9) * (np.pi / 4), 0.0), [1e-300, 2.0 * np.pi]])
    sv, cv = ctypes.c_double(), ctypes.c_double()
    S, C = np.empty_like(x), np.empty_like(x)
    for i, v in enumerate(x):
        f(float(v), ctypes.byref(sv), ctypes.byref(cv))
        S[i], C[i] = sv.value, cv.value
    for got, ref in ((S, np.sin(x)), (C, np.cos(x))):
        err = np.abs(got - ref)
        assert err.max() <= 2.0 ** -53
        big = np.abs(ref) > 1e-3
        assert np.all(err[big] <= np.spacing(np.abs(ref[big])))
    assert S[50000] == 0.0 and C[50000] == 1.0


def test_tabular_grid_branches(O, W):
    """UseTabularGrid in the oracle: the VDW energy of a type with a grid is
    the interpolation (positions wrapped into the grid's unit cell), other
    types stay direct; the Coulomb grid serves every charged type times its
    charge, ignores the scaling, and never an uncharged type; detaching
    restores the direct sums."""
    grid, geom, ubox, uinv, f_and_derivs = _tricubic_grid(11)
    fw = W.c1_framework(n_rep=1)
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    t_grid = s.add_pseudo_atom(148.0, 3.73, 0.0)
    t_direct = s.add_pseudo_atom(148.0, 3.73, 0.0)
    t_q = s.add_pseudo_atom(79.0, 3.05, -0.35)
    rng = np.random.default_rng(11)
    pts = rng.uniform(0.1, 1.9, (10, 3))
    direct = [s.vdw_at(t_grid, p) for p in pts]
    s.use_grid(t_grid, grid, geom, ubox)
    s.use_coulomb_grid(grid, geom, ubox)
    for p, u0 in zip(pts, direct):
        poly = f_and_derivs(*p)[0]
        tol = 1e-5 * max(1.0, np.abs(grid).max())
        assert abs(s.vdw_at(t_grid, p) - poly) <= tol
        # a lattice translation of the grid's unit cell gives the same value
        assert abs(s.vdw_at(t_grid, p + np.array([2.0, -4.0, 6.0])) - s.vdw_at(t_grid, p)) <= 1e-9 * tol
        assert s.vdw_at(t_direct, p) == u0
        assert abs(s.charge_at(t_q, p) - (-0.35) * poly) <= 0.35 * tol
        assert s.charge_at(t_q, p, scaling=0.5) == s.charge_at(t_q, p)
        assert s.charge_at(t_direct, p) == 0.0
    s.use_grid(t_grid, None, None, None)
    s.use_coulomb_grid(None, None, None)
    assert [s.vdw_at(t_grid, p) for p in pts] == direct


def test_surface_area_geometry(O):
    """SurfaceAreaMove's estimate: an isolated atom exposes its whole sphere
    (4 pi d^2 exactly); a second atom at distance D hides the cap whose
    points lie within its own radius, leaving the fraction (1 + c0)/2 with
    c0 = (d^2 + D^2 - d_B^2) / (2 d D)."""
    box, inv, props = O.cell_make(60.0, 60.0, 60.0)
    s = O.System(box, O.BC_RECTANGULAR, [10.0], [10.0], [10.0], [0.0], [50.0], [3.0])
    t = s.add_pseudo_atom(36.0, 3.31)
    f = 2.0 ** (1.0 / 6.0)
    d = f * 0.5 * (3.31 + 3.0)
    area, counts = s.surface_area(t, 1000, seed=3)
    assert counts[0] == 1000 and area == 4.0 * math.pi * d * d
    D = 3.2
    s2 = O.System(box, O.BC_RECTANGULAR, [10.0, 10.0 + D], [10.0, 10.0], [10.0, 10.0], [0.0, 0.0],
                  [50.0, 40.0], [3.0, 2.6])
    t2 = s2.add_pseudo_atom(36.0, 3.31)
    n = 200000
    area2, c2 = s2.surface_area(t2, n, seed=5)
    dA, dB = d, f * 0.5 * (3.31 + 2.6)
    for mine, other, got in ((dA, dB, c2[0]), (dB, dA, c2[1])):
        c0 = (mine * mine + D * D - other * other) / (2.0 * mine * D)
        p = (1.0 + c0) / 2.0
        assert abs(got / n - p) < 5.0 * math.sqrt(p * (1.0 - p) / n)
    expect = (c2[0] / n) * 4 * math.pi * dA * dA + (c2[1] / n) * 4 * math.pi * dB * dB
    assert abs(area2 - expect) <= 1e-12 * expect


def test_surface_area_none_atom(O):
    """An atom of a NONE pseudo-atom carries no sphere and excludes no point
    (src/mc_moves.c:16456, src/framework.c:11676): placed last (so the other
    atoms keep their counter-stream points) and masked, it leaves area and
    counts exactly those of the system without it; interacting, the same
    atom hides part of its neighbour's sphere and adds its own."""
    box, inv, props = O.cell_make(40.0, 40.0, 40.0)
    base = ([10.0, 20.0], [10.0, 10.0], [10.0, 10.0], [0.0, 0.0], [50.0, 40.0], [3.0, 2.6])
    s0 = O.System(box, O.BC_RECTANGULAR, *base)
    t0 = s0.add_pseudo_atom(36.0, 3.31)
    a0, c0 = s0.surface_area(t0, 20000, seed=11)
    x, y, z, q, e, sg = ([v for v in col] for col in base)
    x.append(11.5)  # 1.5 A from atom 0: inside its contact sphere
    y.append(10.0)
    z.append(10.0)
    q.append(0.0)
    e.append(0.0)
    sg.append(1.0)
    s1 = O.System(box, O.BC_RECTANGULAR, x, y, z, q, e, sg)
    t1 = s1.add_pseudo_atom(36.0, 3.31)
    a1, c1 = s1.surface_area(t1, 20000, seed=11, interacts=[1, 1, 0])
    assert a1 == a0 and list(c1[:2]) == list(c0) and c1[2] == 0
    a2, c2 = s1.surface_area(t1, 20000, seed=11)
    assert c2[0] < c0[0] and c2[2] > 0 and a2 != a0
