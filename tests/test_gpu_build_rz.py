"""The engine-backed r-z fly function (the drop-in for build_rz.make_rz_fly_fn
and its fly_fn, ion_gym/physics/build_rz.py:561-777) against the oracle,
ion by ion: every record row of every channel, and the summary (kind, tof,
mz, end point, r_end, n_col), on

* the einzel lens at its full grid (401 x 101), vacuum, FP64, flown to fate,
  plus one ion born inside an electrode (a one-row birth splat, fate 0);
* the ion funnel at its full grid (1201 x 201) with hard-sphere He on the
  shared Philox stream, a 1 us window of 4000 steps, per-mass ions;
* the funnel again with the deck's ccs_fit default in N2, so that each
  ion's cross section is its own (the oracle is handed the same per-ion
  values from ion_gym_amd.physics.ccs).

The bar is the FP64 parity bar: 1e-9 relative per column (against the
column's largest magnitude), identical kinds and collision counts.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from ion_gym_amd.engine import RZ_CHANNELS
from ion_gym_amd.physics import ccs
from ion_gym_amd.physics.flight_spec import (CollisionSpec, FlightSpec,
                                             IntegrationSpec, SourceSpec)

pytestmark = pytest.mark.gpu

REL = 1e-9


def _births(ions):
    return np.stack([ions[k] for k in ("x", "y", "z", "vx", "vy", "vz",
                                       "tob")], axis=1)


def _check(got, o, ions, cols):
    assert len(got) == len(o["x"])
    for i, (tr, s) in enumerate(got):
        nr = int(o["nrec"][i])
        assert tr.shape == (nr, len(cols)), (i, tr.shape, nr)
        ref = o["rec"][i, :nr]
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-12)
        err = np.abs(tr - ref) / scale
        assert np.all(err <= REL), (i, cols[int(np.argmax(err.max(axis=0)))],
                                    float(err.max()))
        assert s["kind"] == int(o["kind"][i]), (i, s["kind"], o["kind"][i])
        assert abs(s["tof"] - o["t"][i]) <= REL * max(1.0, abs(o["t"][i]))
        for k, ok in (("x_end", "x"), ("y_end", "y"), ("z_end", "z")):
            assert abs(s[k] - o[ok][i]) <= REL * max(1.0, abs(o[ok][i])), k
        assert s["r_end"] == pytest.approx(math.hypot(o["y"][i], o["z"][i]),
                                           rel=REL, abs=REL)
        assert s["mz"] == float(ions["m_amu"][i])
        assert s["n_col"] == int(o["ncol"][i])


def test_einzel_fly_fn_matches_oracle_with_a_birth_in_metal():
    from ion_gym_amd import configs
    from ion_gym_amd.physics.build_rz import make_rz_fly_fn
    c = configs.einzel_rz(n_ions=400, seed=0)
    ions = c["ions"]
    n = len(ions["x"])
    births = _births(ions)
    # one more ion, born inside the middle tube's wall (x 16-26, r 5-6 mm)
    births = np.vstack([births, [21.0, 5.5, 0.0, 0.1, 0.0, 0.0, 0.0]])
    spec = FlightSpec(
        source=SourceSpec(n_ions=n + 1, mz_list=[100.0], seed=0),
        integration=IntegrationSpec(dt_ns=1.0, t_max_us=c["t_max"],
                                    rec_every=25,
                                    record_channels=list(RZ_CHANNELS)))
    fly_fn, cols = make_rz_fly_fn(c["model"], births, spec)
    got = fly_fn.batch(range(n + 1))
    rows = int(c["t_max"] / c["dt"] / 25) + 8
    o = O.fly_rz(c["model"].pack(),
                 dict(dt=c["dt"], t_max=c["t_max"], rec_every=25,
                      rec_rows=rows, flags=[1] * 13, rng=O.RNG_PHILOX), ions)
    _check(got[:n], o, ions, cols)
    assert {s["kind"] for _, s in got[:n]} >= {0, 1}   # splats and exits
    tr, s = got[n]
    assert s["kind"] == 0 and s.get("born_in_metal") and s["n_col"] is None
    assert tr.shape == (1, len(cols))
    assert np.array_equal(tr[0, :7], births[n][[6, 0, 1, 2, 3, 4, 5]])
    assert np.all(np.isnan(tr[0, 7:]))
    # fly_fn(i) is batch([i])
    one = fly_fn(7)
    assert np.array_equal(one[0], got[7][0]) and one[1] == got[7][1]


def _funnel_case(sigma_model):
    from ion_gym_amd import configs
    c = configs.funnel_rz(n_ions=600, seed=3)
    ions = dict(c["ions"])
    n = len(ions["x"])
    gas = "He" if sigma_model == "fixed" else "N2"
    col = CollisionSpec(model="hs", gas=gas, T_k=300.0, P_torr=2.0,
                        sigma_model=sigma_model, sigma_m2=2.0e-18)
    spec = FlightSpec(
        source=SourceSpec(n_ions=1, mz_list=list(ions["m_amu"]), seed=3),
        collisions=col,
        integration=IntegrationSpec(
            dt_ns=0.25, t_max_us=1.0, rec_every=200,
            record_channels=["speed", "ke_ev", "e_field", "radius", "n_col",
                             "path_mm", "ke_tint"]))
    ions["sigma_m2"] = np.array([ccs.ion_sigma_m2(col, m, 1)
                                 for m in ions["m_amu"]])
    from ion_gym_amd.physics.collision3d import gas_kernel_scalars
    g = gas_kernel_scalars(gas, T_k=300.0, P_torr=2.0)
    ogas = dict(model=1, **{k: g[k] for k in ("T_k", "P_pa", "c_star",
                                              "c_bar", "sig1d", "m_gas")})
    return c, ions, spec, ogas


@pytest.mark.parametrize("sigma_model", ["fixed", "ccs_fit"])
def test_funnel_fly_fn_matches_oracle_with_hard_sphere(sigma_model):
    from ion_gym_amd.physics.build_rz import make_rz_fly_fn
    c, ions, spec, ogas = _funnel_case(sigma_model)
    n = len(ions["x"])
    fly_fn, cols = make_rz_fly_fn(c["model"], _births(ions), spec)
    got = fly_fn.batch(range(n))
    chans = spec.integration.record_channels
    flags = [int(ch in chans) for ch in RZ_CHANNELS]
    rows = int(1.0 / 0.25e-3 / 200) + 8
    o = O.fly_rz(c["model"].pack(),
                 dict(dt=0.25e-3, t_max=1.0, rec_every=200, rec_rows=rows,
                      flags=flags, gas=ogas, rng=O.RNG_PHILOX), ions)
    _check(got, o, ions, cols)
    assert np.mean([s["n_col"] for _, s in got]) > 50      # collisional
    if sigma_model == "ccs_fit":
        # the fit gives heavier ions a larger cross section, hence more
        # collisions per unit time at equal speed: the cross sections flown
        # are the ions' own, not one value
        assert len(set(np.round(ions["sigma_m2"], 30))) == n
