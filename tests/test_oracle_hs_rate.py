"""Pins of the oracle's hard-sphere collision RATE against kinetic theory.

The oracle restates collision3d.py:265-278 (mean relative speed, mean free
path) and :333-374 (one Bernoulli collision opportunity per step).  Its
thermalisation pin does not depend on the rate, so these checks fix the
rate itself, each against a closed form that shares no expression with the
restatement:

* collision frequency of thermal ions in field-free gas: an ion population
  in equilibrium with the gas meets  n sigma <g>  collisions per unit time,
  <g> = sqrt(8 kT / (pi mu)) the mean relative speed of two Maxwellians
  (mu the reduced mass).  Ion masses 4, 40 and 720 in He separate the
  plausible slips: a dropped x1000 or P in Torr (factors 1000 and 133), the
  gas mean speed c_bar in place of the mean relative speed (sqrt(2) at
  m = 4, 5 % at m = 40) or the ion speed in its place (factor 2.5 at m = 40);
* the mobility gate C1-b of ion_gym's tests/validate_collisions_c1.py:64-82
  (C60, m 720, in He at 1 Torr and 298 K, sigma 1.25e-18 m^2, E = 0.2 V/mm)
  against the first-order Chapman-Enskog (Mason-Schamp) hard-sphere
  mobility  K = 3 q / (16 N) sqrt(2 pi / (mu kT)) / sigma  and against the
  reference's K0 = 4.31 cm^2/(V s), both at the reference's 5 % gate;
* the diffusion gate C1-c (:85-130): field-free D against Einstein's
  kT K / q and the closure D q / (kT K_measured) = 1, at its 10 % gate.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

KB = 1.380649e-23
QE = 1.602176634e-19
AMU = 1.66053906660e-27
TORR = 101325.0 / 760.0
M_HE = 4.002602


def _gas(T, P_pa, m_gas=M_HE):
    """The kernel gas scalars, from their definitions (mm/us)."""
    mg = m_gas * AMU
    return dict(model=1, T_k=T, P_pa=P_pa,
                c_star=math.sqrt(2 * KB * T / mg) / 1000.0,
                c_bar=math.sqrt(8 * KB * T / (math.pi * mg)) / 1000.0,
                sig1d=math.sqrt(KB * T / mg) / 1000.0, m_gas=m_gas)


def _pack(Ez_v_per_m, nx=401, nr=41, h=1.0):
    nu = 2 * nr - 1
    return dict(EzA=np.full((nx, nu), float(Ez_v_per_m)),
                EuA=np.zeros((nx, nu)), EzK=np.zeros((0, nx, nu)),
                EuK=np.zeros((0, nx, nu)), ele=np.zeros((nx, nu)),
                label=np.zeros((nx, nu), np.uint8), u0=-(nr - 1) * h, mm=h,
                ch_kind=np.zeros(0, np.int64), ch_om=np.zeros(0),
                ch_ph=np.zeros(0), ch_duty=np.zeros(0))


def _thermal_ions(n, m_amu, T, sigma, x0, seed):
    """n ions at (x0, 0, 0) with Maxwellian velocities at T (mm/us)."""
    rng = np.random.default_rng(seed)
    s = math.sqrt(KB * T / (m_amu * AMU)) / 1000.0
    v = rng.normal(0.0, s, (n, 3))
    return dict(x=np.full(n, x0), y=np.zeros(n), z=np.zeros(n),
                vx=v[:, 0], vy=v[:, 1], vz=v[:, 2], tob=np.zeros(n),
                m_amu=np.full(n, m_amu),
                acc=np.full(n, QE / (m_amu * AMU)),
                sigma_m2=np.full(n, sigma),
                seed=seed * 100000 + np.arange(n, dtype=np.int64))


def _mu_kg(m_ion, m_gas=M_HE):
    return m_ion * m_gas / (m_ion + m_gas) * AMU


@pytest.mark.parametrize("m_ion", [4.0, 40.0, 720.0])
def test_thermal_collision_frequency_is_n_sigma_mean_relative_speed(m_ion):
    T, P, sigma = 300.0, 1.0 * TORR, 1.0e-18
    n_gas = P / (KB * T)
    g_mean = math.sqrt(8 * KB * T / (math.pi * _mu_kg(m_ion)))  # m/s
    nu = n_gas * sigma * g_mean * 1e-6                          # per us
    # nu dt = 2e-3: a Bernoulli opportunity per step undercounts by
    # about nu dt / 2 = 0.1 %, well inside the 1 % bar
    dt = 2e-3 / nu
    n, per_ion = 400, 500.0
    t_max = per_ion / nu
    ions = _thermal_ions(n, m_ion, T, sigma, 200.0, seed=int(m_ion))
    r = O.fly_rz(_pack(0.0), dict(dt=dt, t_max=t_max, gas=_gas(T, P),
                                  rng=O.RNG_MT), ions)
    assert np.all(r["kind"] == 2)
    total = float(r["ncol"].sum())
    rate = total / float(np.sum(r["steps"] * dt))
    se = 1.0 / math.sqrt(total)               # Poisson counting error
    assert se < 0.0025
    assert rate / nu == pytest.approx(1.0, abs=0.01), (rate, nu)


# ion_gym tests/validate_collisions_c1.py:42-50: C60 in He at 1 Torr, 298 K
C60_M, C60_T, C60_P, C60_SIGMA = 720.0, 298.0, 133.28, 1.25e-18
C60_K0 = 4.31                     # cm^2/(V s), Gillig et al.


def _mason_schamp_K(m_ion, T, P, sigma, m_gas=M_HE):
    """First-order Chapman-Enskog hard-sphere mobility, m^2/(V s)."""
    N = P / (KB * T)
    return (3.0 * QE / (16.0 * N)) * math.sqrt(
        2.0 * math.pi / (_mu_kg(m_ion, m_gas) * KB * T)) / sigma


def _c60_K_mm():
    """K in mm^2/(V us) from the reference's K0 (its :46-48)."""
    K = C60_K0 * (760.0 / 1.0) * (C60_T / 273.15)
    return K * 1e2 / 1e6


def test_mason_schamp_reproduces_the_gillig_mobility():
    # the gate's cross section was derived from K0 = 4.31 by Mason-Schamp:
    # the closed form used below must give the reference's K back
    K = _mason_schamp_K(C60_M, C60_T, C60_P, C60_SIGMA)   # m^2/(V s)
    assert K == pytest.approx(_c60_K_mm(), rel=0.005)


def _c60_flight(E_v_per_mm, n, t_us, seed):
    ions = _thermal_ions(n, C60_M, C60_T, C60_SIGMA, 100.0, seed)
    r = O.fly_rz(_pack(E_v_per_mm * 1e3),
                 dict(dt=4e-4, t_max=t_us, gas=_gas(C60_T, C60_P),
                      rng=O.RNG_MT), ions)
    assert np.all(r["kind"] == 2)
    return ions, r


def _effective_time(t, K):
    """t - tau (1 - exp(-t / tau)) with tau = m K / q (K in mm^2/(V us)):
    the displacement of an ion born with a thermal velocity lags the
    steady drift and diffusion by the momentum relaxation time."""
    tau = C60_M * AMU * K / QE * 1e6           # us
    return t - tau * (1.0 - math.exp(-t / tau))


@pytest.fixture(scope="module")
def c60_mobility():
    """Drift speed (mm/us) of 240 C60 ions over 150 us at 0.2 V/mm
    (the reference's C1-b settings: dt 4e-4 us)."""
    E = 0.2
    ions, r = _c60_flight(E, 240, 150.0, seed=71)
    dx = r["x"] - ions["x"]
    t = float(np.mean(r["t"]))
    vd = float(dx.mean()) / t
    for _ in range(3):
        # thermal starts: the drift lags by the momentum relaxation time
        eff = _effective_time(t, vd / E)
        vd = float(dx.mean()) / eff
    se = float(dx.std()) / eff / math.sqrt(len(dx))
    return E, vd, se


def test_c1b_mobility_matches_mason_schamp_and_gillig(c60_mobility):
    E, vd, se = c60_mobility
    K_ms = _mason_schamp_K(C60_M, C60_T, C60_P, C60_SIGMA)     # = mm^2/(V us)
    assert se / (K_ms * E) < 0.015
    assert vd / (K_ms * E) == pytest.approx(1.0, abs=0.05)
    assert vd / (_c60_K_mm() * E) == pytest.approx(1.0, abs=0.05)


def test_c1c_diffusion_matches_einstein_and_closes(c60_mobility):
    E, vd, _ = c60_mobility
    n, t_us = 1200, 60.0
    ions, r = _c60_flight(0.0, n, t_us, seed=977)
    d2 = ((r["x"] - ions["x"]) ** 2 + (r["y"] - ions["y"]) ** 2
          + (r["z"] - ions["z"]) ** 2)
    # Maxwellian start: <|dr|^2> = 6 D (t - tau (1 - e^{-t/tau})) with
    # tau = m K / q the momentum relaxation time (2.7 us here, 4 % of t)
    t = float(np.mean(r["t"]))
    D = float(np.mean(d2) / (6.0 * _effective_time(t, vd / E)))
    se = math.sqrt(2.0 / (3.0 * n))
    assert se < 0.03
    kT_q = KB * C60_T / QE                                    # V
    D_einstein = kT_q * _c60_K_mm()
    assert D / D_einstein == pytest.approx(1.0, abs=0.10)
    assert D / (kT_q * vd / E) == pytest.approx(1.0, abs=0.10)
