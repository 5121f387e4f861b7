"""Blocked pockets in the oracle (src/framework.c:6168-6300,
src/mc_moves.c:7882-7886), pinned by brute force over periodic images, by
the reference's own pocket files (tests/golden/blocks.json) and by the
consistency of Widom records with and without pockets."""
import json
import os

import numpy as np
import pytest

from helpers import oracle_molecule, oracle_system

HERE = os.path.dirname(os.path.abspath(__file__))


def _blocks():
    with open(os.path.join(HERE, "golden", "blocks.json")) as f:
        return json.load(f)


def test_replication_follows_reference_order(W):
    """Pocket i of unit cell (j, k, l) sits at box @ ((f_i + (j,k,l)) / N), in
    pocket-major order (src/framework.c:6209-6226)."""
    d = _blocks()["DDR"]
    box = W.cell_box(40.0, 40.0, 40.0)
    cen, rad = W.block_pocket_centres(d["fractional_centres"], d["radii"], box, (2, 2, 2))
    assert len(cen) == 15 * 8 and len(rad) == 15 * 8
    f = np.asarray(d["fractional_centres"])
    assert np.allclose(cen[0], f[0] / 2 * 40.0)
    assert np.allclose(cen[1], (f[0] + [0, 0, 1]) / 2 * 40.0)
    assert np.allclose(cen[8], f[1] / 2 * 40.0)
    assert rad[7] == d["radii"][0] and rad[8] == d["radii"][1]


def _brute(cen, rad, ubox, p):
    """Any periodic image of any pocket centre within its radius of p."""
    m = np.asarray(ubox).reshape(3, 3)
    shifts = np.array([[i, j, k] for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)])
    imgs = shifts @ m
    for c, r in zip(cen, rad):
        d = np.linalg.norm(c + imgs - p, axis=1)
        if d.min() < r:
            return True, np.abs(d.min() - r)
    return False, np.min([np.abs(np.linalg.norm(c + imgs - p, axis=1) - r).min()
                          for c, r in zip(cen, rad)])


@pytest.mark.parametrize("which", ["c1", "c2"])
@pytest.mark.parametrize("invert", [False, True])
def test_blocked_pocket_brute_force(W, O, which, invert):
    """For radii below half the smallest perpendicular width the unit-cell
    minimum image finds the only image that can be within reach, so the
    flag equals a brute-force search over images."""
    fw = W.c1_framework() if which == "c1" else W.c2_framework()
    s = oracle_system(fw)
    d = _blocks()["MFI_Olson"]
    nrep = (2, 2, 2) if which == "c1" else (1, 1, 1)
    cen, rad = W.block_pocket_centres(d["fractional_centres"], d["radii"], fw.box, nrep)
    s.set_block_pockets(cen, rad, fw.unit_box, invert=invert)
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (3000, 3)) @ np.asarray(fw.box).reshape(3, 3)
    hits = 0
    for p in pts:
        ref, margin = _brute(cen, rad, fw.unit_box, p)
        if margin < 1e-9:
            continue
        got = s.BlockedPocket(p)
        assert got == (ref != invert), p
        hits += ref
    assert 100 < hits < 2900


def test_widom_records_with_pockets(W, O):
    """On the same trials, an insertion is rejected with stage 5 exactly when
    it was accepted without pockets and an atom of its grown molecule lies in
    a pocket; every other record is unchanged."""
    fw = W.c1_framework(n_rep=1)
    s = oracle_system(fw, charge_method=O.CHARGE_EWALD)
    _, mo = oracle_molecule(s, W.CO2)
    d = _blocks()["DDR"]
    cen, rad = W.block_pocket_centres(d["fractional_centres"], d["radii"], fw.box, (1, 1, 1))
    n = 120
    trials = [O.stream_trials(s, mo, 99, i) for i in range(n)]
    plain = np.array([O.widom_insertion_trials(s, mo, *t) for t in trials])
    s.set_block_pockets(cen, rad, fw.unit_box)
    blocked = np.array([O.widom_insertion_trials(s, mo, *t) for t in trials])
    st = blocked[:, 13]
    for i in range(n):
        if plain[i, 13] != 4:
            assert np.array_equal(plain[i], blocked[i])
            continue
        fb, cord, _ = trials[i]
        sel_fb, sel_or = int(plain[i, 7]), int(plain[i, 8])
        start = fb[sel_fb]
        atoms = start + cord[sel_or]
        expect = any(s.BlockedPocket(a) for a in atoms)
        assert (st[i] == 5) == expect, i
        if expect:
            assert blocked[i, 0] == 0.0 and blocked[i, 6] == 0.0
        else:
            assert np.array_equal(plain[i], blocked[i])
    assert (st == 5).sum() > 0
