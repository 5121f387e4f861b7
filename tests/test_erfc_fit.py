"""The FP32 erfc of the mixed real-space Ewald kernels (erfcx_poly, evaluated
by erfc_fit and erfc_over_r in raspa2_amd/csrc/rhg_kernels.hip).

erfc_fit(x): its coefficients, evaluated as the kernel does in FP32 (FMA
Horner, one exp2 of the rounded -x^2 log2 e), stay within 4e-7 absolute of
scipy's erfc on [0, 5], the range the Ewald cut-offs use (a pair term enters a
site's sum with its absolute error), and within 1.2e-6 relative up to x = 3.2
(alpha rc at the default Ewald precision), 3e-6 on [0, 5].

erfc_over_r(r^2, 1/r): the form of the image-list pair loops, from r^2 and
1/r with t = 1/(1 + (alpha/2 r^2)(1/r)) and exp2(r^2 (-alpha^2 log2 e)): r
times its error against erfc(alpha r)/r (the erfc-equivalent error) stays
within 4e-7 for r up to the 12 A cut-off at alpha = 0.2651 (Ewald precision
1e-6), and the relative error within 1.2e-6."""
import os
import re

import numpy as np
from scipy.special import erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


def _coefficients():
    src = open(os.path.join(ROOT, "raspa2_amd", "csrc", "rhg_kernels.hip")).read()
    body = src[src.index("float erfcx_poly(float t)"):]
    body = body[:body.index("return p;")]
    # the nine coefficients, high to low
    return [float(v) for v in re.findall(r"(-?\d+\.\d+(?:e-?\d+)?)f", body)]


def _fma32(a, b, c):
    return f32(np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64))


def _poly(t):
    hi_to_lo = _coefficients()
    assert len(hi_to_lo) == 9
    p = np.full_like(t, f32(hi_to_lo[0]))
    for c in hi_to_lo[1:]:
        p = _fma32(p, t, np.full_like(t, f32(c)))
    return p


def test_erfc_mixed_accuracy():
    x = np.linspace(0.0, 5.0, 1_000_001).astype(f32)
    t = (f32(1.0) / _fma32(np.full_like(x, 0.5), x, np.ones_like(x))).astype(f32)
    p = _poly(t)
    x2 = (x * x).astype(f32)
    yh = (x2 * f32(-1.44269502)).astype(f32)
    e = np.exp2(yh.astype(np.float64)).astype(f32)
    got = (p * e).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 4e-7
    rel = np.abs(got - ref) / ref
    assert np.max(rel[x <= f32(3.2)]) <= 1.2e-6
    err = np.max(rel)
    assert err <= 3e-6, err


def test_erfc_over_r_accuracy():
    alpha = 0.2651
    a = f32(alpha)
    bh, ke = f32(f32(0.5) * a), f32(-f32(a * a) * f32(1.44269502))
    r = np.linspace(0.5, 12.0, 1_000_001)
    r2 = (r * r).astype(f32)
    irs = (1.0 / np.sqrt(r2.astype(np.float64))).astype(f32)
    t = (f32(1.0) / _fma32(f32(bh * r2), irs, np.ones_like(r2))).astype(f32)
    e = np.exp2(f32(r2 * ke).astype(np.float64)).astype(f32)
    got = f32(f32(_poly(t) * e) * irs).astype(np.float64)
    rr = np.sqrt(r2.astype(np.float64))
    ref = erfc(alpha * rr) / rr
    assert np.max(np.abs(got - ref) * rr) <= 4e-7
    assert np.max(np.abs(got - ref) / ref) <= 1.2e-6
