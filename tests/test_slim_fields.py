"""The SLIM board's bases: the parallel-board Laplace solver against the
closed form of a single Fourier mode, a field-free uniform board, and the
pack's channel layout (11 sets, square TW phases, sin RF) and mirror
symmetry about the mid-gap plane."""
import math

import numpy as np
import pytest

from ion_gym_amd import configs


def test_single_mode_matches_the_closed_form():
    h, Px, Pz, gap, ny = 0.1, 64, 32, 2.0, 21
    X = np.arange(Px) * h
    m = 3
    kx = 2 * math.pi * m / (Px * h)
    B = np.cos(kx * X)[:, None] * np.ones(Pz)[None, :]
    Ex, Ey, Ez = configs._slab_fields(B, h, gap, ny)
    y = np.arange(ny) * h - 0.5 * gap
    C = np.cosh(kx * y) / math.cosh(kx * gap / 2)
    S = np.sinh(kx * y) / math.cosh(kx * gap / 2)
    # phi = cos(kx x) C(y): Ex = kx sin(kx x) C, Ey = -cos(kx x) kx S
    want_x = kx * np.sin(kx * X)[None, :, None] * C[:, None, None]
    want_y = -kx * np.cos(kx * X)[None, :, None] * S[:, None, None]
    assert np.allclose(Ex, want_x, atol=1e-12)
    assert np.allclose(Ey, want_y, atol=1e-12)
    assert np.abs(Ez).max() < 1e-12


def test_uniform_board_is_field_free():
    B = np.full((32, 64), 7.0)
    Ex, Ey, Ez = configs._slab_fields(B, 0.1, 1.0, 11)
    for E in (Ex, Ey, Ez):
        assert np.abs(E).max() < 1e-12


@pytest.fixture(scope="module")
def slim():
    return configs.slim_3d(coarse=4)


def test_slim_pack_layout(slim):
    pk = slim["pack"]
    nx, ny, nz = slim["shape"]
    assert pk["EAx"].shape == (nx, ny, nz)
    assert pk["ExK"].shape == (10, nx, ny, nz)
    assert list(pk["ch_kind"]) == [2] * 8 + [0, 0]
    assert np.allclose(pk["ch_ph"][8:], [0.0, math.pi])
    # the 8 TW phases step by 2 pi / 8
    assert np.allclose(np.diff(pk["ch_ph"][:8]), -2 * math.pi / 8)


def test_slim_fields_mirror_about_mid_gap(slim):
    """Both boards carry the same pattern: Ex, Ez even and Ey odd in
    y - gap/2."""
    pk = slim["pack"]
    for k in (0, 3, 8):
        assert np.allclose(pk["ExK"][k], pk["ExK"][k][:, ::-1, :], atol=1e-9)
        assert np.allclose(pk["EzK"][k], pk["EzK"][k][:, ::-1, :], atol=1e-9)
        assert np.allclose(pk["EyK"][k], -pk["EyK"][k][:, ::-1, :],
                           atol=1e-9)


def test_slim_requests_span_the_declared_ranges(slim):
    rng = np.random.default_rng(0)
    sizes, amps, freqs = [], [], []
    for _ in range(200):
        ions, tune = configs.slim_request(slim, rng, t_flight_us=1.0)
        sizes.append(len(ions["x"]))
        amps.append(2 * tune["ch_amp"][0])
        v = tune["ch_om"][0] / (2 * math.pi) * 8 * slim["pitch_mm"] * 1e3
        freqs.append(v)
        assert tune["ch_off"][0] == tune["ch_amp"][0]
        assert 200 <= ions["m_amu"][0] <= 3000
    assert 1e3 <= min(sizes) and max(sizes) <= 5e5
    assert 10 <= min(amps) and max(amps) <= 40
    assert 50 <= min(freqs) and max(freqs) <= 300
