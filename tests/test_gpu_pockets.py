"""Blocked pockets on the GPU against the oracle (src/framework.c:6246,
src/mc_moves.c:7882-7886): flags of BlockedPocket and Widom accept/reject
records bit-exact on identical trials."""
import json
import os

import numpy as np
import pytest

from helpers import gpu_molecule, oracle_molecule, oracle_system, same_bits

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _pockets(W, fw, name, nrep):
    with open(os.path.join(HERE, "golden", "blocks.json")) as f:
        d = json.load(f)[name]
    return W.block_pocket_centres(d["fractional_centres"], d["radii"], fw.box, nrep)


@pytest.mark.parametrize("which,invert", [("c1", False), ("c2", False), ("c2", True)])
def test_blocked_pocket_flags(W, O, R, which, invert):
    fw = W.c1_framework() if which == "c1" else W.c2_framework()
    cen, rad = _pockets(W, fw, "MFI_Olson", (2, 2, 2) if which == "c1" else (1, 1, 1))
    s = oracle_system(fw)
    s.set_block_pockets(cen, rad, fw.unit_box, invert=invert)
    ctx = fw.context()
    ctx.set_block_pockets(cen, rad, fw.unit_box, invert=invert)
    rng = np.random.default_rng(8)
    pts = rng.uniform(-0.2, 1.2, (20000, 3)) @ np.asarray(fw.box).reshape(3, 3)
    got = ctx.BlockedPocket(pts)
    ref = np.array([s.BlockedPocket(p) for p in pts])
    assert np.array_equal(got, ref)
    assert 0 < ref.sum() < len(ref)


@pytest.mark.parametrize("which", ["c0_methane", "c1_co2"])
def test_widom_with_pockets(W, O, R, which):
    """Stage-5 rejections (and every other flag) equal the oracle's on the
    same counter-stream trials; accepted W to rounding."""
    if which == "c0_methane":
        fw, probe, ch, n = W.c0_framework(), W.METHANE, O.CHARGE_NONE, 3000
        cen, rad = _pockets(W, fw, "DDR", (1, 1, 1))
    else:
        fw, probe, ch, n = W.c1_framework(), W.CO2, O.CHARGE_EWALD, 300
        cen, rad = _pockets(W, fw, "DDR", (2, 2, 2))
    s = oracle_system(fw, charge_method=ch)
    _, mo = oracle_molecule(s, probe)
    s.set_block_pockets(cen, rad, fw.unit_box)
    ctx = fw.context(charge_method=ch)
    _, mg = gpu_molecule(ctx, probe)
    ctx.set_block_pockets(cen, rad, fw.unit_box)
    ro = O.widom_run_stream(s, mo, 21, 0, n)
    _, rg = ctx.widom_run(mg, 21, n, records=True)
    for f in (6, 7, 8, 9, 10, 13):
        assert np.array_equal(ro[:, f], rg[:, f]), f
    assert (ro[:, 13] == 5).sum() > 0
    acc = ro[:, 13] == 4
    assert same_bits(ro[acc, 2], rg[acc, 2])
    assert np.max(np.abs(ro[acc, 0] - rg[acc, 0]) / ro[acc, 0]) <= 1e-9
