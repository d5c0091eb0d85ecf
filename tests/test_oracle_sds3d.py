"""The oracle's 3-D SDS flight (tracer3d._fly3d_sds) pinned without the
reference: with diffusion off, the damped kick in a uniform field has the
closed form v_n = v_d + (v_0 - v_d) rho^n, rho = 1 - factor gamma dt,
v_d = v_gas + a / gamma, and x_n = x_0 + dt (v_1 + ... + v_n); with
diffusion on and no field, <r^2> grows as 6 D t (the mean free path
normalisation of the jump table); with both, the mean drift is K E."""
import math

import numpy as np

from oracle import oracle as O


def _cell(n, e_vpmm, mfp_on=True, coarse=2, **kw):
    from ion_gym_amd import configs
    c = configs.drift_3d(n_ions=n, coarse=coarse, e_vpmm=e_vpmm, **kw)
    ions = dict(c["ions"])
    if not mfp_on:
        sds = ions["sds"].copy()
        sds[:, 1] = 0.0
        ions["sds"] = sds
    return c, ions


def test_damped_kick_closed_form_without_diffusion():
    c, ions = _cell(50, 5.0, mfp_on=False)
    n_steps = 3000
    fl = dict(dt=c["dt"], t_max=1e9, max_steps=n_steps, gas=c["gas"],
              rng=O.RNG_PHILOX)
    o = O.fly_xyz(c["pack"], fl, ions)
    assert np.all(o["kind"] == -2) and np.all(o["steps"] == n_steps)
    g = ions["sds"][0, 0]
    dt = c["dt"]
    factor = (1.0 - math.exp(-g * dt)) / (g * dt)
    rho = 1.0 - factor * g * dt
    a = ions["acc"] * 1e-6 * 5.0
    vd = np.array([np.zeros(50), np.zeros(50), a / g])
    v0 = np.array([ions["vx"], ions["vy"], ions["vz"]])
    x0 = np.array([ions["x"], ions["y"], ions["z"]])
    vn = vd + (v0 - vd) * rho ** n_steps
    # x_n - x_0 = dt sum_{k=1..n} v_k
    geo = rho * (1.0 - rho ** n_steps) / (1.0 - rho)
    xn = x0 + dt * (n_steps * vd + (v0 - vd) * geo)
    for k, key in enumerate(("vx", "vy", "vz")):
        assert np.allclose(o[key], vn[k], rtol=1e-9, atol=1e-15), key
    for k, key in enumerate(("x", "y", "z")):
        assert np.allclose(o[key], xn[k], rtol=1e-10, atol=1e-12), key
    assert np.allclose(o["t"], ions["tob"] + n_steps * dt, rtol=1e-12)


def test_diffusion_and_drift():
    # no field: the spread grows as 6 D t
    c, ions = _cell(3000, 0.0)
    n_steps = 4000
    fl = dict(dt=c["dt"], t_max=1e9, max_steps=n_steps, gas=c["gas"],
              rng=O.RNG_PHILOX)
    o = O.fly_xyz(c["pack"], fl, ions)
    alive = o["kind"] == -2
    assert alive.mean() > 0.95
    d2 = ((o["x"] - ions["x"]) ** 2 + (o["y"] - ions["y"]) ** 2
          + (o["z"] - ions["z"]) ** 2)[alive]
    t = n_steps * c["dt"]
    assert abs(d2.mean() / (6.0 * c["D"] * t) - 1.0) < 0.06
    # 5 V/mm: the mean axial speed is K E
    c, ions = _cell(3000, 5.0)
    o = O.fly_xyz(c["pack"], dict(fl, gas=c["gas"]), ions)
    alive = o["kind"] == -2
    vz = (o["z"] - ions["z"])[alive] / t
    assert abs(vz.mean() / (c["K_mm2_per_V_us"] * 5.0) - 1.0) < 0.02


def test_transporter_unwraps_to_the_straight_line():
    """Field-free 3-D RK4 flight along z through a transporter (accept
    8.005 mm, emit 2 mm, direction +1, three passes): the pass count in
    each record row is the number of accept crossings of the unwrapped
    path z_u = z0 + v t, the flight ends with fate 4 on the third pass,
    and z + 2 (accept - emit) there is the straight line (the last pass
    is not shifted)."""
    nx, ny, nz, h = 5, 5, 101, 0.1
    pack = dict(EAx=np.zeros((nx, ny, nz)), EAy=np.zeros((nx, ny, nz)),
                EAz=np.zeros((nx, ny, nz)), ExK=np.zeros((0, nx, ny, nz)),
                EyK=np.zeros((0, nx, ny, nz)), EzK=np.zeros((0, nx, ny, nz)),
                ele=np.zeros((nx, ny, nz), bool), h_mm=h,
                ch_kind=np.zeros(0, np.int64), ch_om=np.zeros(0),
                ch_ph=np.zeros(0), ch_amp=np.zeros(0), ch_off=np.zeros(0),
                ch_duty=np.zeros(0))
    n = 3
    ions = dict(x=np.full(n, 0.2), y=np.full(n, 0.2), z=np.array([1.0, 1.3, 4.0]),
                vx=np.zeros(n), vy=np.zeros(n), vz=np.array([1.0, 0.7, 1.3]),
                tob=np.zeros(n), m_amu=np.full(n, 100.0),
                acc=np.full(n, 9.6e5), sigma_m2=np.zeros(n),
                seed=np.arange(n, dtype=np.int64))
    tp = dict(axis="z", accept_mm=8.005, emit_mm=2.0, direction=1.0,
              max_passes=3)
    dt = 1e-8
    o = O.fly_xyz(pack, dict(dt=dt, t_max=1e-3, transporter=tp, rec_every=37,
                             rec_rows=2000), ions)
    d = tp["accept_mm"] - tp["emit_mm"]
    assert np.all(o["kind"] == 4)
    for i in range(n):
        rows = o["rec"][i, :o["nrec"][i]]
        t_us = rows[:, 0]
        zu = ions["z"][i] + ions["vz"][i] * t_us
        passes = np.floor((zu - tp["accept_mm"]) / d) + 1
        passes = np.maximum(passes, 0)
        # the last row is the third pass, recorded unshifted
        assert np.array_equal(rows[:-1, 10], passes[:-1])
        assert rows[-1, 10] == 3
        assert np.allclose(rows[:-1, 3] + rows[:-1, 10] * d, zu[:-1],
                           rtol=0, atol=1e-9)
        assert abs(o["z"][i] + 2 * d - (ions["z"][i] + ions["vz"][i]
                                         * (o["t"][i]))) < 1e-9
