"""Fits the FP32 complementary error function used by the mixed-precision
real-space Ewald kernels and prints the coefficients for rhg_common.h.

    erfc(x) = exp(-x^2) * P(t),   t = 1 / (1 + x/2),   0 <= x <= 5

P (degree 8, monomials in t) approximates the scaled function
erfcx(x) = exp(x^2) erfc(x), which is smooth and between 0.11 and 1 on the
interval; the fit is least squares in the relative error, reweighted
iteratively towards the minimax solution.  In exact arithmetic the relative
error is 3.5e-8; evaluated in FP32 as the kernel does (t rounded, Horner
with FMA, the exponent -x^2 log2(e) carried as head + tail) it is below
5e-7 on [0, 5] (tests/test_erfc_fit.py), about the error that rounding x to
FP32 already causes in erfc(x) near x = 3 (d ln erfc / dx = -2x).

    python raspa2_amd/csrc/fit_erfc.py
"""
import numpy as np
from scipy.special import erfcx

P_T = 0.5   # t = 1 / (1 + P_T x)
DEG = 8
X_MAX = 5.0


def fit():
    x = np.linspace(0.0, X_MAX, 400001)
    t = 1.0 / (1.0 + P_T * x)
    y = erfcx(x)
    V = np.vander(t, DEG + 1, increasing=True)
    w = 1.0 / y
    wts = np.ones_like(x)
    for _ in range(40):
        A = V * (w * wts)[:, None]
        c, *_ = np.linalg.lstsq(A, y * w * wts, rcond=None)
        r = np.abs(V @ c - y) / y
        wts = wts * (1.0 + r / r.max())
    return c, r.max()


if __name__ == "__main__":
    c, err = fit()
    print(f"// max relative error (exact arithmetic) {err:.3e}")
    for k, v in enumerate(c):
        print(f"  {float(np.float32(v)):.9g}f, // t^{k}")
