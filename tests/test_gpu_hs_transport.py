"""Hard-sphere transport on the GPU, both precisions, against the oracle
flying the reference's algorithm and MT19937 stream and against kinetic
theory, at the reference's C-1 gate conditions (ion_gym
tests/validate_collisions_c1.py:42-130: C60, m 720, in He at 1 Torr and
298 K, sigma 1.25e-18 m^2, dt 4e-4 us):

* mobility at E = 0.2 V/mm: the steady drift speed (the mean displacement
  over t - tau (1 - e^{-t/tau}), tau = m K / q, for thermal starts) of
  200,000 ions over 150 us
  within 1 % of the oracle's (6,000 ions; sampling error 0.2 %) and within
  2 % of the first-order Chapman-Enskog (Mason-Schamp) hard-sphere
  mobility, which the reference's gate holds to 5 %;
* diffusion at zero field: D from the mean squared displacement of
  200,000 ions over 60 us (Maxwellian start, corrected for the momentum
  relaxation time m K / q) within 1 % of Einstein's kT K / q with the
  drift test's own measured K (the fluctuation-dissipation closure, exact
  at zero field; sampling error 0.2 %), and within 1 % or four standard
  errors of the oracle's D (2,000 ions).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KB = 1.380649e-23
QE = 1.602176634e-19
AMU = 1.66053906660e-27
M, T, P, SIGMA, M_HE = 720.0, 298.0, 133.28, 1.25e-18, 4.002602
DT = 4e-4


def _gas():
    mg = M_HE * AMU
    return dict(model=1, T_k=T, P_pa=P,
                c_star=math.sqrt(2 * KB * T / mg) / 1000.0,
                c_bar=math.sqrt(8 * KB * T / (math.pi * mg)) / 1000.0,
                sig1d=math.sqrt(KB * T / mg) / 1000.0, m_gas=M_HE,
                flow=(0.0, 0.0, 0.0))


def _pack(E_vpmm, nx=401, nr=41, h=1.0):
    nu = 2 * nr - 1
    return dict(EzA=np.full((nx, nu), E_vpmm * 1e3), EuA=np.zeros((nx, nu)),
                EzK=np.zeros((0, nx, nu)), EuK=np.zeros((0, nx, nu)),
                ele=np.zeros((nx, nu)), label=np.zeros((nx, nu), np.uint8),
                u0=-(nr - 1) * h, mm=h, ch_kind=np.zeros(0, np.int64),
                ch_om=np.zeros(0), ch_ph=np.zeros(0), ch_duty=np.zeros(0))


def _ions(n, seed):
    rng = np.random.default_rng(seed)
    s = math.sqrt(KB * T / (M * AMU)) / 1000.0
    v = rng.normal(0.0, s, (n, 3))
    return dict(x=np.full(n, 100.0), y=np.zeros(n), z=np.zeros(n),
                vx=v[:, 0], vy=v[:, 1], vz=v[:, 2], tob=np.zeros(n),
                m_amu=np.full(n, M), acc=np.full(n, QE / (M * AMU)),
                sigma_m2=np.full(n, SIGMA),
                seed=seed * 1000000 + np.arange(n, dtype=np.int64))


def _gpu(engine, E, ions, t_us, prec):
    from ion_gym_amd.engine import Flight
    fp = engine.upload_rz(_pack(E), prec)
    g = engine.fly_batch(fp, Flight(DT, t_us, gas=_gas()), ions)
    fp.release()
    assert np.all(g["fate"] == 2)
    return g["x"], g["y"], g["z"], g["t_us"]


def _oracle(E, ions, t_us):
    r = O.fly_rz(_pack(E), dict(dt=DT, t_max=t_us, gas=_gas(),
                                rng=O.RNG_MT), ions)
    assert np.all(r["kind"] == 2)
    return r["x"], r["y"], r["z"], r["t"]


def _mason_schamp_K():
    """First-order hard-sphere mobility, m^2/(V s) = mm^2/(V us)."""
    N = P / (KB * T)
    mu = M * M_HE / (M + M_HE) * AMU
    return 3 * QE / (16 * N) * math.sqrt(2 * math.pi / (mu * KB * T)) / SIGMA


def _effective_time(t, K):
    """t - tau (1 - exp(-t / tau)), tau = m K / q: the drift and the
    diffusion of ions born with thermal velocities lag the steady state by
    the momentum relaxation time (2.6 us here)."""
    tau = M * AMU * K / QE * 1e6                    # us
    return t - tau * (1.0 - math.exp(-t / tau))


def _drift(x, x0, t, E):
    """Steady drift speed (mm/us) and its standard error."""
    tm = float(np.mean(t))
    dx = x - x0
    vd = float(dx.mean()) / tm
    for _ in range(3):                  # tau depends on K = vd / E
        eff = _effective_time(tm, vd / E)
        vd = float(dx.mean()) / eff
    return vd, float(dx.std()) / eff / math.sqrt(len(dx))


def _diffusion(x, y, z, ions, t, K):
    d2 = (x - ions["x"]) ** 2 + (y - ions["y"]) ** 2 + (z - ions["z"]) ** 2
    D = d2 / (6.0 * _effective_time(float(np.mean(t)), K))
    return D.mean(), D.std() / math.sqrt(len(D))


@pytest.fixture(scope="module")
def oracle_runs():
    E = 0.2
    io = _ions(6000, 71)
    x, y, z, t = _oracle(E, io, 150.0)
    vd_o, se_vd_o = _drift(x, io["x"], t, E)
    iz = _ions(2000, 977)
    x, y, z, t = _oracle(0.0, iz, 60.0)
    D_o, se_D_o = _diffusion(x, y, z, iz, t, vd_o / E)
    return dict(vd=vd_o, se_vd=se_vd_o, D=D_o, se_D=se_D_o)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hs_mobility_and_diffusion(engine, oracle_runs, prec):
    o = oracle_runs
    E = 0.2
    io = _ions(200_000, 171)
    x, y, z, t = _gpu(engine, E, io, 150.0, prec)
    vd, se_vd = _drift(x, io["x"], t, E)
    assert math.hypot(se_vd, o["se_vd"]) / vd < 0.0025
    assert vd / o["vd"] == pytest.approx(1.0, abs=0.01), (vd, o["vd"])
    assert vd / (_mason_schamp_K() * E) == pytest.approx(1.0, abs=0.02)
    K = vd / E
    iz = _ions(200_000, 1977)
    x, y, z, t = _gpu(engine, 0.0, iz, 60.0, prec)
    D, se_D = _diffusion(x, y, z, iz, t, K)
    assert se_D / D < 0.0025
    einstein = KB * T / QE * K
    assert D / einstein == pytest.approx(1.0, abs=0.01), (D, einstein)
    tol = max(0.01 * o["D"], 4.0 * math.hypot(se_D, o["se_D"]))
    assert abs(D - o["D"]) <= tol, (D, o["D"], se_D, o["se_D"])
