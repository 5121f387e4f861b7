"""Pins of the oracle's finite-loading terms: guest-guest energies
(src/inter_energy.c:321-520), the Fourier Delta with stored adsorbate sums
and the Old-molecule branch (src/ewald.c:6859-7565), and the adsorbate terms
of the CBMC/Widom bookkeeping (src/cbmc.c:623-795,1154-1478,2762-2849,
3056-3185; src/mc_moves.c:7844-7933).  Each check is chosen so that a
dropped term, a swapped sign or a mixed-up molecule fails it."""
import math

import numpy as np
import pytest
from scipy.special import erf, erfc

from helpers import oracle_system, random_positions


def _co2(c, v):
    v = np.asarray(v, float)
    v /= np.linalg.norm(v)
    c = np.asarray(c, float)
    return np.array([c + 1.16 * v, c, c - 1.16 * v])


def _loaded(O, W, n_mol=6, seed=3, ion=False):
    """The 25.8 A charged cell with n_mol CO2 molecules (and optionally a
    +0.5 e single-site ion as the last molecule, and then a framework net
    charge of 0.576 e so that every UIon term is exercised)."""
    fw = W.c1_framework(n_rep=1)
    if ion:
        fw.charge = fw.charge + 0.001
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    types = [s.add_pseudo_atom(*site) for site in W.CO2.sites]
    t_ion = s.add_pseudo_atom(50.0, 3.0, 0.5)
    rng = np.random.default_rng(seed)
    mols = [_co2(p, rng.normal(size=3)) for p in random_positions(fw.box, n_mol, seed)]
    start, tys, xyz = [0], [], []
    for m in mols:
        tys += types
        xyz += list(m)
        start.append(start[-1] + 3)
    if ion:
        tys.append(t_ion)
        xyz.append(random_positions(fw.box, 1, seed + 1)[0])
        start.append(start[-1] + 1)
    s.set_adsorbates(start, tys, np.array(xyz))
    return fw, s, types, t_ion, start, tys, np.array(xyz)


def _lj(eps_a, sig_a, eps_b, sig_b, rr, rc2, K2E):
    e = math.sqrt(eps_a * K2E * eps_b * K2E)
    sg = 0.5 * (sig_a + sig_b)
    u = lambda r2: 4.0 * e * ((sg * sg / r2) ** 6 - (sg * sg / r2) ** 3)
    return u(rr) - u(rc2)


def test_guest_guest_brute_force(O, W):
    """Guest-guest VDW and real-space Coulomb at a position: LB-mixed shifted
    LJ and C qa qb erfc(alpha r)/r over the minimum images of every atom of
    every molecule but the excluded one."""
    fw, s, types, t_ion, start, tys, xyz = _loaded(O, W)
    units = O.units()
    C, K2E, alpha = units["coulomb"], 1.0 / units["energy_to_kelvin"], s.info()["alpha"]
    L = fw.box[0]
    eps = {t: e for t, (e, _, _) in zip(types, W.CO2.sites)}
    sig = {t: sg for t, (_, sg, _) in zip(types, W.CO2.sites)}
    chg = {t: q for t, (_, _, q) in zip(types, W.CO2.sites)}
    for p in random_positions(fw.box, 8, 29):
        for ta in types:
            for excl in (-1, 2):
                ev, ec = 0.0, 0.0
                for m in range(len(start) - 1):
                    if m == excl:
                        continue
                    for a in range(start[m], start[m + 1]):
                        d = p - xyz[a]
                        d -= L * np.round(d / L)
                        rr = float(d @ d)
                        if rr < 144.0:
                            ev += _lj(eps[ta], sig[ta], eps[tys[a]], sig[tys[a]], rr, 144.0, K2E)
                            r = math.sqrt(rr)
                            ec += C * chg[ta] * chg[tys[a]] * erfc(alpha * r) / r
                gv = s.inter_vdw_at(ta, p, excl)
                gc = s.inter_charge_at(ta, p, excl)
                assert abs(gv - ev) <= 1e-11 * max(1.0, abs(ev))
                assert abs(gc - ec) <= 1e-11 * max(1.0, abs(ec))


def test_guest_guest_pair_symmetry(O, W):
    """The interaction of molecule A with the rest seen from A equals the sum
    of what each other molecule sees of A (pair symmetry of mixing and
    minimum image), and excluding a molecule removes exactly its share."""
    fw, s, types, t_ion, start, tys, xyz = _loaded(O, W, n_mol=4, seed=5)
    nm = len(start) - 1

    def seen_by(mol, exclude):
        ev = ec = 0.0
        for a in range(start[mol], start[mol + 1]):
            ev += s.inter_vdw_at(tys[a], xyz[a], exclude)
            ec += s.inter_charge_at(tys[a], xyz[a], exclude)
        return ev, ec

    for A in range(nm):
        va, ca = seen_by(A, A)
        vb = cb = 0.0
        for B in range(nm):
            if B == A:
                continue
            # what B sees of A alone = B's view with everything but A minus B's view with all but A and B:
            # computed directly by adsorbate sets holding only A
            sub = [A]
            st, ty, xy = [0], [], []
            for m in sub:
                ty += tys[start[m]:start[m + 1]]
                xy += list(xyz[start[m]:start[m + 1]])
                st.append(st[-1] + start[m + 1] - start[m])
            s.set_adsorbates(st, ty, np.array(xy))
            v, c = seen_by_other(s, tys, xyz, start, B)
            vb += v
            cb += c
        s.set_adsorbates(start, tys, xyz)
        assert abs(va - vb) <= 1e-10 * max(1.0, abs(va))
        assert abs(ca - cb) <= 1e-10 * max(1.0, abs(ca))


def seen_by_other(s, tys, xyz, start, B):
    ev = ec = 0.0
    for a in range(start[B], start[B + 1]):
        ev += s.inter_vdw_at(tys[a], xyz[a], -1)
        ec += s.inter_charge_at(tys[a], xyz[a], -1)
    return ev, ec


def _frame(O, fw, extra_xyz=(), extra_q=()):
    """Framework plus extra charges as one framework: its Fourier energy
    sum_k kf |S(k)|^2 and net charge."""
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    ex = np.asarray(extra_xyz, float).reshape(-1, 3)
    qq = np.r_[q, np.asarray(extra_q, float)]
    n = len(ex)
    s2 = O.System(box, bc, np.r_[x, ex[:, 0]], np.r_[y, ex[:, 1]], np.r_[z, ex[:, 2]], qq,
                  np.r_[eps, np.ones(n)], np.r_[sig, np.ones(n)],
                  charge_method=O.CHARGE_EWALD, ewald_precision=1e-6)
    return s2.fourier_self_framework(), qq.sum()


@pytest.mark.parametrize("ion", [False, True])
def test_fourier_move_is_energy_difference(O, W, ion):
    """Insertion, deletion and reinsertion: host + adsorbate Delta + self +
    excluded equals the change of sum_k kf |S_f + S_ads|^2 plus
    UIon (Q_total^2 change); the host part alone equals the change of the
    framework-adsorbate cross sum."""
    fw, s, types, t_ion, start, tys, xyz = _loaded(O, W, n_mol=5, seed=7, ion=ion)
    qmap = {t: q for t, (_, _, q) in zip(types, W.CO2.sites)}
    qmap[t_ion] = 0.5
    q_all = np.array([qmap[t] for t in tys])
    uion = s.info()["uion"]
    F0, Q0 = _frame(O, fw, xyz, q_all)
    Ff, Qf = _frame(O, fw)
    assert (abs(Qf) > 0.5) == ion

    def cross(xx, qq):
        Fx, _ = _frame(O, fw, xx, qq)
        Fonly = _frame_only(O, fw, xx, qq)
        return Fx - Ff - Fonly

    H0 = cross(xyz, q_all)
    new = _co2([9.0, 11.0, 4.0], [0.2, 0.9, -0.3])
    last = len(start) - 2
    moves = {
        "insert": (types, new, -1, np.r_[xyz, new], np.r_[q_all, [qmap[t] for t in types]]),
        "delete": ((), None, 1, np.r_[xyz[:3], xyz[6:]], np.r_[q_all[:3], q_all[6:]]),
        "reinsert": (types, new, 0, np.r_[new, xyz[3:]], np.r_[[qmap[t] for t in types], q_all[3:]]),
    }
    if ion:
        p_ion = xyz[-1] + np.array([3.0, -2.0, 1.0])
        moves["ion_delete"] = ((), None, last, xyz[:-1], q_all[:-1])
        moves["ion_insert"] = ([t_ion], p_ion[None, :], -1, np.r_[xyz, p_ion[None, :]], np.r_[q_all, [0.5]])
    for name, (nt, nx, old, xyz1, q1) in moves.items():
        out = s.fourier_move(nt, nx, old)
        F1, Q1 = _frame(O, fw, xyz1, q1)
        want = (F1 - F0) + uion * (Q1 * Q1 - Q0 * Q0)
        got = out[0] + out[1] + out[2] + out[3]
        assert abs(got - want) <= 1e-8 * max(1.0, abs(F0)), (name, got, want)
        H1 = cross(xyz1, q1)
        want_h = (H1 - H0) + 2.0 * uion * Qf * (Q1 - Qf - (Q0 - Qf))
        assert abs(out[0] - want_h) <= 1e-8 * max(1.0, abs(H0)), (name, out[0], want_h)


def _frame_only(O, fw, xx, qq):
    """sum_k kf |S_X|^2 of the extra charges alone (the framework's atoms kept
    with zero charge so the k table is the same)."""
    box, bc, x, y, z, q, eps, sig, mass = fw.oracle_args()
    ex = np.asarray(xx, float).reshape(-1, 3)
    n = len(ex)
    s2 = O.System(box, bc, np.r_[x, ex[:, 0]], np.r_[y, ex[:, 1]], np.r_[z, ex[:, 2]],
                  np.r_[np.zeros(len(x)), qq], np.r_[eps, np.ones(n)], np.r_[sig, np.ones(n)],
                  charge_method=O.CHARGE_EWALD, ewald_precision=1e-6)
    return s2.fourier_self_framework()


def test_widom_ghost_adsorbates_change_nothing(O, W):
    """Adsorbates with no LJ well and no charge leave every Widom record
    bit-identical."""
    fw = W.c1_framework(n_rep=1)
    s0 = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    s1 = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    mols = []
    for s in (s0, s1):
        ty = [s.add_pseudo_atom(*site) for site in W.CO2.sites]
        ghost = s.add_pseudo_atom(0.0, 3.0, 0.0)
        mols.append(O.molecule(ty, W.CO2.body, start=W.CO2.start, k_fb=4, k_or=6))
    pts = random_positions(fw.box, 12, 31)
    s1.set_adsorbates(list(range(13)), [ghost] * 12, pts)
    r0 = O.widom_run_stream(s0, mols[0], 9, 0, 60)
    r1 = O.widom_run_stream(s1, mols[1], 9, 0, 60)
    assert np.array_equal(r0, r1)
    # RASPA2's driver: the same Rosenbluth sums, and an ideal-gas sum of
    # Volume / (N + 1) per insertion with the N = 12 (ghost) molecules present
    # (src/mc_moves.c:7861-7863) against Volume / 1 for the empty framework
    n0, b0 = O.widom_run_mt(s0, mols[0], 17, 6, 10)
    n1, b1 = O.widom_run_mt(s1, mols[1], 17, 6, 10)
    V = s0.info()["volume"]
    assert n0 == n1 == 60 and np.array_equal(b0[:, :3], b1[:, :3])
    # (sums of equal terms: equal to the product up to rounding)
    assert np.allclose(b0[:, 3], b0[:, 0] * V, rtol=1e-13, atol=0)
    assert np.allclose(b1[:, 3], b1[:, 0] * (V / 13.0), rtol=1e-13, atol=0)


def test_widom_adsorbate_bookkeeping(O, W):
    """With host-supplied trials the records of a loaded system decompose
    into their parts: U_ads_vdw and U_ads_real are the guest-guest energies
    of the grown molecule (first bead, then the placed beads), the Fourier
    parts equal fourier_move of the grown molecule, and W = W1 W2
    exp(-beta (U_real + U_ads_real)) exp(-beta (UF_host + UF_ads))."""
    fw, s, types, t_ion, start, tys, xyz = _loaded(O, W, n_mol=8, seed=11)
    mol = O.molecule(types, W.CO2.body, start=W.CO2.start, k_fb=5, k_or=8)
    beta = s.info()["beta"]
    checked = 0
    for i in range(40):
        fb, cord, usel = O.stream_trials(s, mol, 41, i)
        rec = O.widom_insertion_trials(s, mol, fb, cord, usel)
        if rec[13] != 4:
            continue
        checked += 1
        sf, so = int(rec[7]), int(rec[8])
        first = fb[sf]
        pos = [first + cord[so][j] for j in range(3)]
        pos[W.CO2.start] = first
        st = W.CO2.start
        ev = s.inter_vdw_at(types[st], first, 8)
        ec = s.inter_charge_at(types[st], first, 8)
        pv = pc = 0.0
        for j in range(3):
            if j != st:
                pv += s.inter_vdw_at(types[j], pos[j], 8)
                pc += s.inter_charge_at(types[j], pos[j], 8)
        assert rec[14] == ev + pv and rec[15] == ec + pc
        fo = s.fourier_move(types, np.array(pos), -1)
        assert rec[4] == fo[0] and rec[5] == fo[1]
        w = rec[11] * rec[12] * math.exp(-beta * (rec[3] + rec[15] + 0.0))
        w *= math.exp(-beta * (fo[0] + fo[1] + 0.0))
        assert abs(rec[0] - w) <= 1e-12 * w
        du = rec[14] + rec[2] + rec[15] + rec[3] + rec[4] + rec[5]
        assert rec[1] == du
    assert checked >= 10


def test_retrace_matches_insertion(O, W):
    """CBMC detailed balance at finite loading: insert a molecule with given
    trials, add it to the adsorbates, then retrace it with the same trial set
    (its selected trials moved to slot 0): the retrace weight equals the
    insertion weight without its Fourier factor, W_ins = W_ret
    exp(-beta (UF_host + UF_ads)), with the same energies; without excluding
    itself the retrace overlaps."""
    fw, s, types, t_ion, start, tys, xyz = _loaded(O, W, n_mol=10, seed=17)
    mol = O.molecule(types, W.CO2.body, start=W.CO2.start, k_fb=5, k_or=8)
    beta = s.info()["beta"]
    st = W.CO2.start
    done = 0
    for i in range(60):
        fb, cord, usel = O.stream_trials(s, mol, 43, i)
        rec = O.widom_insertion_trials(s, mol, fb, cord, usel)
        if rec[13] != 4:
            continue
        sf, so = int(rec[7]), int(rec[8])
        old = np.array([fb[sf] + cord[so][j] for j in range(3)])
        old[st] = fb[sf]
        fo = s.fourier_move(types, old, -1)
        # the molecule joins the adsorbates as molecule number 10
        s.set_adsorbates(start + [start[-1] + 3], tys + list(types), np.r_[xyz, old])
        fb_r = np.r_[fb[sf:sf + 1], np.delete(fb, sf, axis=0)]
        cord_r = np.r_[cord[so:so + 1], np.delete(cord, so, axis=0)]
        ret = O.widom_retrace_trials(s, mol, 10, old, fb_r, cord_r)
        assert ret[13] == 4 and ret[7] == 0 and ret[8] == 0
        for f in (2, 3, 14, 15):  # host and guest-guest energies of the configuration
            assert ret[f] == rec[f]
        assert abs(ret[11] - rec[11]) <= 1e-14 * rec[11]
        assert abs(ret[12] - rec[12]) <= 1e-14 * rec[12]
        w_ins = ret[0] * math.exp(-beta * (fo[0] + fo[1] + 0.0))
        assert abs(w_ins - rec[0]) <= 1e-12 * rec[0]
        assert ret[4] == 0.0 and ret[5] == 0.0
        bad = O.widom_retrace_trials(s, mol, -1, old, fb_r, cord_r)
        assert int(bad[9]) & 1  # trial 0 sits on its own atoms
        s.set_adsorbates(start, tys, xyz)
        done += 1
    assert done >= 10
