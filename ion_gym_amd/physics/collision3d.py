"""Gas scalars of the hard-sphere model.

Restates the host half of ion_gym/physics/collision3d.py: the gas mass
presets (:59-60), gas_mass (:161-168), the vacuum placeholders (:66-67) and
gas_kernel_scalars (:207-245), which turn a gas declaration into the
scalars every flight kernel takes.  The per-step physics itself runs in the
engine's kernels (ion_gym_amd/csrc/ig_device.h).
"""
import math

from .constants import K_B as KB, KG_AMU, TORR_PA

GASES = {"He": 4.002602, "H2": 2.01588, "N2": 28.0134, "air": 28.9647,
         "Ar": 39.948, "CO2": 44.0095, "Kr": 83.798, "Xe": 131.293}
_VACUUM_REF_T_K = 298.0
_VACUUM_REF_M_GAS = 4.0
MAX_PARTNER_TRIES = 10_000
HS_DT_PER_COLLISION_TIME = 20.0


def gas_mass(gas):
    """Preset name -> amu, or an explicit mass (collision3d.py:161-168)."""
    if isinstance(gas, str):
        if gas not in GASES:
            raise KeyError(f"unknown gas '{gas}'; presets: "
                           f"{sorted(GASES)} -- or pass a mass in amu")
        return GASES[gas]
    return float(gas)


def gas_kernel_scalars(gas=None, T_k=None, P_torr=None, P_pa=None):
    """(enabled, T_k, P_pa, c_star, c_bar, sig1d, m_gas) for a gas given by
    name/mass, temperature and pressure; ``gas=None`` is vacuum: P = 0 with
    the inert thermal placeholders (collision3d.py:230-245)."""
    if gas is None:
        T, mg, P = _VACUUM_REF_T_K, _VACUUM_REF_M_GAS, 0.0
        enabled = False
    else:
        T = float(T_k)
        mg = gas_mass(gas)
        P = float(P_pa) if P_pa is not None else float(P_torr) * TORR_PA
        if not (T > 0.0 and P > 0.0):
            raise ValueError(f"invalid gas: T = {T_k!r} K, P = {P!r} Pa "
                             f"must be > 0")
        enabled = True
    return dict(
        enabled=enabled, T_k=T, P_pa=P,
        c_star=math.sqrt(2 * KB * T / (mg * KG_AMU)) / 1000.0,
        c_bar=math.sqrt(8 * KB * T / (math.pi * mg * KG_AMU)) / 1000.0,
        sig1d=math.sqrt(KB * T / (mg * KG_AMU)) / 1000.0,
        m_gas=mg)


def collision_time_ns(gas, T_k, P_torr, sigma_m2):
    """Mean collision time of an ion at rest, 1/(n sigma c_bar_gas), in ns
    (the number collision3d.hs_step_table reports, :114-126)."""
    T = float(T_k)
    n_gas = float(P_torr) * TORR_PA / (KB * T)
    m_gas = gas_mass(gas) * KG_AMU
    c_bar = math.sqrt(8.0 * KB * T / (math.pi * m_gas))
    return 1.0e9 / (n_gas * sigma_m2 * c_bar)
