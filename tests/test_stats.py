"""Ensemble statistics against hand-computed answers: the reference's own
synthetic ensemble (tests/test_stats.py:52-95 of ion_gym: fate counts,
transmission, per-mass resolving power with population sigma), the
transmitted-fate rule and mass blocks, and landing groups by declared
plane with the plane's normal."""
import math

import numpy as np
import pytest

from ion_gym_amd.physics.flight_spec import BoundsSpec, FlightSpec
from ion_gym_amd.physics.stats import (FWHM_PER_SIGMA,
                                       auto_transmitted_fate, compute_stats,
                                       enabled_planes, mz_of_results)


class _R:
    def __init__(self, s):
        self.summary = s


def test_synthetic_ensemble():
    res, mz = [], []
    for t in (10.0, 10.1, 9.9, 10.0):
        res.append(_R(dict(kind=3, tof=t, x_end=30.0, y_end=0.1, z_end=0.0,
                           r_end=0.1, n_col=0)))
        mz.append(100.0)
    for t in (20.0, 20.2):
        res.append(_R(dict(kind=3, tof=t, x_end=30.0, y_end=-0.1,
                           z_end=0.0, r_end=0.1, n_col=0)))
        mz.append(400.0)
    for k in (0, 0, 1, 2):
        res.append(_R(dict(kind=k, tof=1.0, x_end=5.0, y_end=2.0, z_end=0.0,
                           r_end=2.0, n_col=0)))
        mz.append(100.0)
    st = compute_stats(res, transmitted_fate=3, mz=mz)
    assert st.n_total == 10 and st.fate_counts == {3: 6, 0: 2, 1: 1, 2: 1}
    assert abs(st.transmission - 0.6) < 1e-12
    assert [g.mz for g in st.per_mz] == [100.0, 400.0]
    xs = [10.0, 10.1, 9.9, 10.0]
    mu = sum(xs) / 4
    sd = math.sqrt(sum((x - mu) ** 2 for x in xs) / 4)
    r_want = mu / (2 * FWHM_PER_SIGMA * sd)
    g = st.per_mz[0]
    assert abs(g.tof_mean - mu) < 1e-12
    assert abs(g.resolving_power - r_want) / r_want < 1e-12
    assert st.impact_spread["x"][0] == 5.0 and st.n_transmitted == 6
    # the landing group of fate 3: normal x (no spread), centroid (y, z)
    ps = st.landing[3]
    assert ps.normal == "x" and ps.n == 6
    # the summaries carry no mz, so the group counts as one mass
    assert ps.tof_n == 6
    both = [dict(r.summary, mz=m) for r, m in zip(res, mz)]
    ps = compute_stats(both, transmitted_fate=3).landing[3]
    assert ps.tof_n == 0 and ps.tof_note.startswith("mixed masses")
    with pytest.raises(ValueError, match="transmitted_fate is required"):
        compute_stats(res)
    with pytest.raises(ValueError, match="not a fate code"):
        compute_stats(res, transmitted_fate=9)


def test_auto_fate_mass_blocks_and_planes():
    spec = FlightSpec()
    assert auto_transmitted_fate(spec) == 1
    spec.bounds = BoundsSpec(x_max_on=True, x_max=30.0, z_min_on=True,
                             z_min=-4.0)
    assert auto_transmitted_fate(spec) == 3
    spec.source.mz_list = [100.0, 400.0]
    spec.source.n_ions = 2
    assert mz_of_results(spec, 5) == [100.0, 100.0, 400.0, 400.0, 100.0]
    planes = enabled_planes(spec)
    assert planes == [("x=+30", "x", 30.0), ("z=-4", "z", -4.0)]
    rng = np.random.default_rng(0)
    res = []
    for i in range(40):        # hits on x = 30 with a beam spot in (y, z)
        res.append(dict(kind=3, tof=5.0 + 0.01 * i, mz=100.0, x_end=30.0,
                        y_end=rng.normal(0, 0.2), z_end=rng.normal(0, 0.3),
                        vx_end=1.0, vy_end=0.01, vz_end=0.0, ke_end=5.0))
    for i in range(10):        # hits on z = -4
        res.append(dict(kind=3, tof=3.0, mz=100.0, x_end=rng.uniform(0, 30),
                        y_end=0.0, z_end=-4.0, vx_end=0.0, vy_end=0.0,
                        vz_end=-1.0))
    st = compute_stats(res, transmitted_fate=3, planes=planes)
    a, b = st.landing["3:x=+30"], st.landing["3:z=-4"]
    assert a.n == 40 and a.normal == "x" and a.in_plane == ("y", "z")
    assert b.n == 10 and b.normal == "z" and b.plane_label == "z=-4"
    ys = np.array([r["y_end"] for r in res[:40]])
    assert abs(a.sigma[0] - ys.std()) < 1e-12
    assert abs(a.div_mrad_rms - 10.0) < 0.01 and a.ke_n == 40
    assert a.tof_n == 40 and b.tof_note == "n < 2 — no width"
