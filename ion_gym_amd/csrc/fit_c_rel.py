"""Coefficients of the FP32-basis mode's mean-relative-speed polynomials.

c_rel(sp) / c_bar = (s + 1/(2 s)) (sqrt(pi)/2) erf(s) + exp(-s^2) / 2 with
s = sp / c* (ion_gym/physics/collision3d.py:265-272) is an even, entire
function of s, i.e. an entire function G(u) of u = s^2.  Two least-squares
fits in a Chebyshev basis, converted to monomials which the kernel
evaluates by Horner's rule in FP32 (ig_device.h):
* s in [0, 1) (where heavy ions in a light gas nearly always are): degree
  5 in y = 2 u - 1 on 200,001 points (c_rel_small_f);
* s in [1, 4): degree 12 in x = u / 8 - 1, fitted on all of [0, 4] on
  400,001 points (c_rel_mid_f).
Above s = 4 the kernel uses (sqrt(pi)/2)(s + 1/(2 s)), whose relative
error there is below 3e-11.  Run to print the coefficients; the CPU test
tests/test_fp32_helpers.py checks the header's values against the exact
formula.
"""
import numpy as np
from scipy.special import erf

DEGREE = 12
DEGREE_SMALL = 5


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


def fit_small(degree=DEGREE_SMALL):
    u = np.linspace(0.0, 1.0, 200001)
    y = 2.0 * u - 1.0
    c = np.polynomial.chebyshev.chebfit(y, exact(np.sqrt(u)), degree)
    return np.polynomial.chebyshev.cheb2poly(c)


if __name__ == "__main__":
    print("c_rel_small_f (s < 1), monomials in y = 2 s^2 - 1:")
    for k, a in enumerate(fit_small()):
        print(f"  {a:.9e}f,  // y^{k}")
    print("c_rel_mid_f (1 <= s < 4), monomials in x = s^2 / 8 - 1:")
    for k, a in enumerate(fit()):
        print(f"  {a:.9e}f,  // x^{k}")
