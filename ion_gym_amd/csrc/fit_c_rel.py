"""Coefficients of the FP32-basis mode's mean-relative-speed polynomial.

c_rel(sp) / c_bar = (s + 1/(2 s)) (sqrt(pi)/2) erf(s) + exp(-s^2) / 2 with
s = sp / c* (ion_gym/physics/collision3d.py:265-272) is an even, entire
function of s, i.e. an entire function G(u) of u = s^2.  On s in [0, 4]
it is fitted by least squares in the Chebyshev basis of x = u / 8 - 1 on
400,001 points (degree 12) and converted to monomials in x, which the
kernel evaluates by Horner's rule in FP32 (ig_device.h, c_rel_ratio_f);
above s = 4 the kernel uses (sqrt(pi)/2)(s + 1/(2 s)), whose relative
error there is below 3e-11.  Run to print the coefficients; the CPU test
tests/test_fp32_helpers.py checks the header's values against the exact
formula.
"""
import numpy as np
from scipy.special import erf

DEGREE = 12


def exact(s):
    s = np.asarray(s, float)
    out = np.ones_like(s)
    m = s > 0
    out[m] = ((s[m] + 0.5 / s[m]) * 0.5 * np.sqrt(np.pi) * erf(s[m])
              + 0.5 * np.exp(-s[m] ** 2))
    return out


def fit(degree=DEGREE):
    u = np.linspace(0.0, 16.0, 400001)
    x = u / 8.0 - 1.0
    c = np.polynomial.chebyshev.chebfit(x, exact(np.sqrt(u)), degree)
    return np.polynomial.chebyshev.cheb2poly(c)


if __name__ == "__main__":
    for k, a in enumerate(fit()):
        print(f"  {a:.9e}f,  // x^{k}")
