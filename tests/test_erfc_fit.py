"""The FP32 erfc of the mixed real-space Ewald kernels (erfc_fit, called by
erfc_mixed, in raspa2_amd/csrc/rhg_kernels.hip): its coefficients, evaluated as the kernel
does in FP32 (FMA Horner, one exp2 of the rounded -x^2 log2 e), stay within
4e-7 absolute of scipy's erfc on [0, 5], the range the Ewald cut-offs use
(a pair term enters a site's sum with its absolute error), and within 1.2e-6
relative up to x = 3.2 (alpha rc at the default Ewald precision), 3e-6 on
[0, 5]."""
import os
import re

import numpy as np
from scipy.special import erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _coefficients():
    src = open(os.path.join(ROOT, "raspa2_amd", "csrc", "rhg_kernels.hip")).read()
    body = src[src.index("float erfc_fit(float x)"):]
    body = body[:body.index("// exp(-x^2)")]
    vals = [float(v) for v in re.findall(r"(-?\d+\.\d+(?:e-?\d+)?)f", body)]
    # 0.5f and 1.0f (t), then the nine coefficients high to low
    return vals[2:11]


def _fma32(a, b, c):
    return np.float32(a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64))


def test_erfc_mixed_accuracy():
    hi_to_lo = _coefficients()
    assert len(hi_to_lo) == 9
    x = np.linspace(0.0, 5.0, 1_000_001).astype(np.float32)
    t = (np.float32(1.0) / _fma32(np.full_like(x, 0.5), x, np.ones_like(x))).astype(np.float32)
    p = np.full_like(x, np.float32(hi_to_lo[0]))
    for c in hi_to_lo[1:]:
        p = _fma32(p, t, np.full_like(x, np.float32(c)))
    f = np.float32
    x2 = (x * x).astype(f)
    yh = (x2 * f(-1.44269502)).astype(f)
    e = np.exp2(yh.astype(np.float64)).astype(f)
    got = (p * e).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 4e-7
    rel = np.abs(got - ref) / ref
    assert np.max(rel[x <= np.float32(3.2)]) <= 1.2e-6
    err = np.max(rel)
    assert err <= 3e-6, err
