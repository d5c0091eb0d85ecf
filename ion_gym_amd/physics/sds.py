"""Host side of the SDS (statistical diffusion simulation) gas model.

Restates the per-ion numbers of ion_gym/physics/sds.py: the jump-table
reader (:91-131), the kernels' jump-length draw (_icdf_steps_at, :134-151,
vectorised here for the host's mean-square integral), the mean squared jump
by the midpoint rule (:172-191) and transport_params (:360-407) for an ion
whose reduced mobility K0 is declared in the deck's gas.  The per-step jump
runs in the engine's kernels.

The jump table.  The reference ships sds_jump_icdf.dat, a Monte Carlo
(100,000 hard-sphere walks of 100,000 collisions per mass ratio) of the
distance from the start, tabulated as 1002 quantiles.  It is not vendored
here.  ``load_diffusion_statistics`` reads that file format, so a user with
ion_gym installed flies the reference's exact table;
``default_statistics`` supplies the limit the Monte Carlo estimates: after
N = 1e5 collisions the displacement is a sum of N weakly correlated steps,
an isotropic 3-D Gaussian by the central limit theorem, so its length
follows the chi distribution with 3 degrees of freedom (the Maxwell law).
The kernels normalise every draw by the table's own mean square
(mfp_mm = 6 D N / (V <d^2>)), so only the shape enters the flight.
"""
import math
import os

import numpy as np

from .constants import ATM_TORR, E_CHG, K_B, KG_AMU, T_STP_K

N_DIST_COLLISIONS = 100000
N_DIST = 5
N_DIST_POINTS = 1002
_MEAN_SQ_POINTS = 200000


def load_diffusion_statistics(path):
    """(5, 1002) table from a file in the reference's format: numbers
    separated by commas or spaces, ';' starting a comment
    (sds.py:91-121)."""
    out = []
    with open(path, encoding="utf-8") as fh:
        for ln, line in enumerate(fh, 1):
            for tok in line.split(";")[0].replace(",", " ").split():
                try:
                    out.append(float(tok))
                except ValueError:
                    raise ValueError(f"{path}:{ln}: {tok!r} is not a "
                                     f"number") from None
    if len(out) != N_DIST * N_DIST_POINTS:
        raise ValueError(f"{path}: {len(out)} numbers, expected "
                         f"{N_DIST * N_DIST_POINTS}")
    return np.array(out, np.float64).reshape(N_DIST, N_DIST_POINTS)


_DEFAULT = None


def default_statistics():
    """The chi-3 (Maxwell) limit of the jump-length quantiles, in the table
    layout: point i is the quantile at probability i/1000 (i = 0..1000)
    and point 1001 repeats the maximum, the layout of
    tools/regen_sds_tables.py:207-212 in the reference.  The sample maximum
    of 1e5 walks is taken at probability 1 - 1e-5.  The same shape is used
    for every mass ratio; the scale sqrt(N) is immaterial (see the module
    note).  ION_GYM_SDS_TABLE names a reference-format file to use
    instead."""
    global _DEFAULT
    if _DEFAULT is None:
        path = os.environ.get("ION_GYM_SDS_TABLE")
        if path:
            _DEFAULT = load_diffusion_statistics(path)
        else:
            from scipy.stats import chi
            q = np.concatenate([[0.0], np.arange(1, 1000) / 1000.0,
                                [1.0 - 1e-5, 1.0 - 1e-5]])
            row = chi(3).ppf(q) * math.sqrt(N_DIST_COLLISIONS)
            _DEFAULT = np.tile(row, (N_DIST, 1))
    return _DEFAULT


def icdf_steps_at(stats, log_mr, u):
    """Vectorised _icdf_steps_at (sds.py:134-151): linear in u within a
    row, log-log between the two rows bracketing the mass ratio."""
    u = np.asarray(u, np.float64)
    x = u * (N_DIST_POINTS - 2)
    i = x.astype(np.int64)
    w = x - i
    k = int(math.floor(log_mr))
    k = min(max(k, 0), N_DIST - 2)
    t = log_mr - k
    a, b = stats[k], stats[k + 1]
    lo = a[i] + w * (a[i + 1] - a[i])
    hi = b[i] + w * (b[i + 1] - b[i])
    llo = np.log10(lo)
    return 10.0 ** (llo + t * (np.log10(hi) - llo))


def mean_sq_jump_steps(log_mr, stats=None):
    """<d^2> of the kernels' draw, midpoint rule over 200,000 points
    (sds.py:172-191)."""
    st = default_statistics() if stats is None else stats
    u = (np.arange(_MEAN_SQ_POINTS) + 0.5) / _MEAN_SQ_POINTS
    d = icdf_steps_at(st, round(float(log_mr), 12), u)
    return float(np.mean(d * d))


def transport_params(mass_amu, charge, gas_mass_amu, T_k, P_torr, K0,
                     stats=None):
    """Per-ion SDS numbers (sds.py:360-407) for an ion of declared reduced
    mobility K0 (cm^2/(V s), per unit charge, in the deck's gas)."""
    m = float(mass_amu)
    if not (m > 0.0):
        raise ValueError(f"SDS: ion mass must be positive, got {mass_amu!r}")
    if not (P_torr > 0.0 and T_k > 0.0):
        raise ValueError("SDS: T and P must be positive")
    K1 = K0 * 1.0e-4 * (ATM_TORR / P_torr) * (T_k / T_STP_K)
    m_kg = m * KG_AMU
    D = K1 * K_B * T_k / E_CHG
    V = math.sqrt(8.0 * K_B * T_k / (math.pi * m_kg)) * 1.0e-3
    log_mr = math.log10(m / float(gas_mass_amu))
    msq = mean_sq_jump_steps(log_mr, stats)
    return dict(damping=E_CHG / (m_kg * K1) * 1.0e-6,
                mfp_mm=6.0 * D * N_DIST_COLLISIONS / (V * msq),
                V_mm_us=V, log_mr_ratio=log_mr, K0=K0,
                K_mm2_per_V_us=abs(int(charge)) * K1, D_mm2_per_us=D)
