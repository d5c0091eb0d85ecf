"""The FP32-basis mode's polynomial helpers, read from the kernel header
and evaluated in FP32 on the host: the fitted exact-Maxwell mean relative
speed (collision3d.py:265-272) within 6e-7 relative over s in [0, 30]
(2.5e-7 below s = 1)."""
import os
import re

import numpy as np
from scipy.special import erf

HDR = os.path.join(os.path.dirname(__file__), "..", "ion_gym_amd", "csrc",
                   "ig_device.h")


def _body(name):
    src = open(HDR).read()
    i = src.index(name)
    j = src.index("\n}\n", i)
    return src[i:j]


def _coeffs(body):
    """The Horner sequence: the initial q value, then each fmaf addend."""
    q0 = re.search(r"float q = (-?[0-9.e+-]+)f;", body).group(1)
    adds = re.findall(r"__builtin_fmaf\(q(?: \* r2)?, (?:x|y|r2|r), "
                      r"(-?[0-9.e+-]+)f\)", body)
    return [float(q0)] + [float(a) for a in adds]


def _exact_crel(s):
    s = np.asarray(s, float)
    out = np.ones_like(s)
    m = s > 0
    out[m] = ((s[m] + 0.5 / s[m]) * 0.5 * np.sqrt(np.pi) * erf(s[m])
              + 0.5 * np.exp(-s[m] ** 2))
    return out


def _horner(c, x):
    f32 = np.float32
    q = np.full_like(x, f32(c[0]))
    for a in c[1:]:
        q = (q * x + f32(a)).astype(np.float32)
    return q


def test_c_rel_polynomials():
    """Both fits and their dispatch (s < 1, 1 <= s < 4, the asymptote)."""
    small = _coeffs(_body("c_rel_small_f(float s)"))
    mid = _coeffs(_body("c_rel_mid_f(float s)"))
    assert len(small) == 6 and len(mid) == 13   # highest degree first
    s = np.linspace(0.0, 30.0, 300001).astype(np.float32)
    f32 = np.float32
    out = np.empty_like(s)
    lo = s < 1.0
    y = (f32(2.0) * s[lo] * s[lo] - f32(1.0)).astype(np.float32)
    out[lo] = _horner(small, y)
    md = (s >= 1.0) & (s < 4.0)
    x = (s[md] * s[md] * f32(0.125) - f32(1.0)).astype(np.float32)
    out[md] = _horner(mid, x)
    hi = s >= 4.0
    out[hi] = f32(0.886226925452758) * (s[hi] + f32(0.5) / s[hi])
    ex = _exact_crel(s.astype(float))
    rel = np.abs(out.astype(float) - ex) / ex
    assert rel[lo].max() < 2.5e-7, rel[lo].max()
    assert rel.max() < 6e-7, rel.max()
