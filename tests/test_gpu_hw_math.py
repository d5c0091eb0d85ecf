"""The hardware FP32 transcendentals of the FP32-basis kernels against the
host's FP64 functions, over the arguments the kernels give them (2^22
evenly spaced points): v_sin_f32 / v_cos_f32 on angles in [-1/2, 1)
revolutions (the reduced drive phases, the random azimuths) within 1.5e-7
absolute, v_log_f32 on (0, 1] (the uniforms of the Exp(1) and normal
draws) within 1.5e-7 relative where |log2 v| > 0.1 and 2.5e-6 absolute
everywhere."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hardware_sin_cos_log():
    path = os.path.join(ROOT, "ion_gym_amd", "libig_probe.so")
    L = ctypes.CDLL(path)
    n = 1 << 22
    out = np.empty((5, n), np.float32)
    print(f"hardware sin/cos/log2 on {n} points", flush=True)
    assert L.igp_hw_math(ctypes.c_int32(n),
                         ctypes.c_void_p(out.ctypes.data)) == 0
    u, s, c, v, lg = (out[k].astype(float) for k in range(5))
    assert u.min() >= -0.5 and u.max() < 1.0
    assert np.max(np.abs(s - np.sin(2 * np.pi * u))) < 1.5e-7
    assert np.max(np.abs(c - np.cos(2 * np.pi * u))) < 1.5e-7
    assert v.min() > 0.0 and v.max() == 1.0
    ref = np.log2(v)
    err = np.abs(lg - ref)
    big = np.abs(ref) > 0.1
    assert np.max(err[big] / np.abs(ref[big])) < 1.5e-7
    assert np.max(err) < 2.5e-6
