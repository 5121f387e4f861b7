"""Host-only parts of the screening service: the admission footprint
estimate (no GPU needed)."""
import numpy as np


def test_footprint_estimate_monotone(W, R, O):
    from raspa2_amd import serve as S
    j = W.c3_job(3)
    a = S.estimate_footprint(j.framework, j.probe, chunk=1 << 12)
    b = S.estimate_footprint(j.framework, j.probe, chunk=1 << 16)
    assert 0 < a < b
    # the per-insertion part is linear in the chunk
    c = S.estimate_footprint(j.framework, j.probe, chunk=1 << 20)
    assert abs((c - b) / ((1 << 20) - (1 << 16)) - (b - a) / ((1 << 16) - (1 << 12))) < 1.0
    # Ewald tables: at least 24 B per k-vector of the oracle's count
    none = S.estimate_footprint(j.framework, j.probe, charge_method=R.CHARGE_NONE, chunk=1 << 12)
    fw = j.framework
    so = O.System(fw.box, fw.bc, fw.xyz[:, 0], fw.xyz[:, 1], fw.xyz[:, 2], fw.charge, fw.eps_k,
                  fw.sigma, mass=fw.mass, charge_method=O.CHARGE_EWALD, ewald_precision=1e-6)
    nk = so.info()["nk"]
    assert a - none >= 24 * nk
