"""Host side of the SDS (statistical diffusion simulation) gas model.

Restates the per-ion numbers of ion_gym/physics/sds.py: the jump-table
reader (:91-131), the kernels' jump-length draw (_icdf_steps_at, :134-151,
vectorised here for the host's mean-square integral), the mean squared jump
by the midpoint rule (:172-191), the paper's estimators of an ion's size
and mobility from its mass (:194-231, Appelhans & Dahl, Int. J. Mass
Spectrom. 244 (2005) 1, Sec. 2.1-2.2), the declared species values of
``collisions.sds_species`` (:234-356) and transport_params /
sds_ion_params (:360-447).  The per-step jump runs in the engine's kernels.

The jump table.  ion_gym ships sds_jump_icdf.dat, a Monte Carlo (100,000
hard-sphere walks of 100,000 collisions per mass ratio 10^0..10^4) of the
distance from the start, tabulated as 1002 quantiles.  ``default_statistics``
reads that file: the one named by ION_GYM_SDS_TABLE, else the one beside an
installed ion_gym package (found without importing it).  Without either it
falls back, with an advisory, to the chi-3 (Maxwell) law, the limit the
Monte Carlo estimates: after 1e5 collisions the displacement is an
isotropic Gaussian sum of weakly correlated steps.  The kernels normalise
every draw by the table's own mean square (mfp_mm = 6 D N / (V <d^2>)), so
only the shape enters a flight; normalised so, the chi-3 law agrees with
the reference's rows 0-2 (mass ratios 1-100) to 0.7 % at every quantile
from 1 % to 99.9 % and departs from rows 3-4 (mass ratios 1e3, 1e4) by up
to 2.0 % and 13 % in the tails (tests/test_sds_table.py states the error
per row).
"""
import importlib.util
import math
import os
import warnings

import numpy as np

from .ccs import IonGymAdvisory
from .constants import ATM_TORR, E_CHG, K_B, KG_AMU, T_STP_K

JUMP_ICDF_FILE = "sds_jump_icdf.dat"
N_DIST_COLLISIONS = 100000
N_DIST = 5
N_DIST_POINTS = 1002
_MEAN_SQ_POINTS = 200000

# paper Sec. 2.1-2.2 (sds.py:58-62): air as the reference gas, the
# constant-density diameter of an ion of mass m, and the two log-log cubic
# fits between the air mobility and the diameter
AIR_MASS_AMU = 28.954515
AIR_DIAMETER_NM = 0.366
MASS_TO_DIAMETER_NM = 0.120415405
_LOGD_OF_LOGK9 = (3.0367, -0.8504, 0.1137, -0.0135)
_LOGK9_OF_LOGD = (4.9137, -1.4491, -0.2772, 0.0717)


def load_diffusion_statistics(path):
    """(5, 1002) table from a file in the reference's format: numbers
    separated by commas or spaces, ';' starting a comment (sds.py:91-121).
    The table is read by position, so a token that is not a number is
    refused with its file, line and text."""
    vals = []
    with open(path, encoding="utf-8") as fh:
        for ln, line in enumerate(fh, 1):
            body = line.split(";", 1)[0]
            for tok in body.replace(",", " ").split():
                try:
                    vals.append(float(tok))
                except ValueError:
                    raise ValueError(
                        f"{path}:{ln}: {tok!r} is not a number. The SDS table "
                        f"is read positionally, so skipping it would shift "
                        f"every later value; fix or remove the token."
                    ) from None
    want = N_DIST * N_DIST_POINTS
    if len(vals) != want:
        raise ValueError(f"{path}: {len(vals)} numbers, expected {want} "
                         f"({N_DIST} rows of {N_DIST_POINTS})")
    return np.asarray(vals, np.float64).reshape(N_DIST, N_DIST_POINTS)


def chi3_statistics():
    """The chi-3 limit of the jump-length quantiles in the table layout:
    point i is the quantile at probability i/1000 (i = 0..1000) and point
    1001 repeats the maximum, taken at probability 1 - 1e-5 (the sample
    maximum of 1e5 walks); the same shape for every mass ratio."""
    from scipy.stats import chi
    q = np.concatenate([[0.0], np.arange(1, 1000) / 1000.0,
                        [1.0 - 1e-5, 1.0 - 1e-5]])
    row = chi(3).ppf(q) * math.sqrt(N_DIST_COLLISIONS)
    return np.tile(row, (N_DIST, 1))


def installed_table_path():
    """sds_jump_icdf.dat of an importable ion_gym package, or None.  The
    package is located, not imported."""
    try:
        spec = importlib.util.find_spec("ion_gym")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    for base in spec.submodule_search_locations:
        p = os.path.join(base, "physics", JUMP_ICDF_FILE)
        if os.path.isfile(p):
            return p
    return None


_DEFAULT = None


def default_statistics():
    """The jump statistics every SDS flight reads, once per process:
    ION_GYM_SDS_TABLE, else the installed ion_gym's table, else the chi-3
    limit (with an IonGymAdvisory saying so)."""
    global _DEFAULT
    if _DEFAULT is None:
        path = os.environ.get("ION_GYM_SDS_TABLE") or installed_table_path()
        if path:
            _DEFAULT = load_diffusion_statistics(path)
        else:
            warnings.warn(
                "[SDS] no sds_jump_icdf.dat found (set ION_GYM_SDS_TABLE or "
                "install ion_gym): SDS jumps use the chi-3 limit of the "
                "tabulated walks, within 0.7 % of ion_gym's table for "
                "ion/gas mass ratios up to 100 and up to 13 % off in the "
                "tails at 1e4", IonGymAdvisory, stacklevel=2)
            _DEFAULT = chi3_statistics()
    return _DEFAULT


def icdf_steps_at(stats, log_mr, u):
    """Vectorised _icdf_steps_at (sds.py:134-151): linear in u within a
    row, log-log between the two rows bracketing the mass ratio
    (extrapolated beyond rows 0 and 4)."""
    u = np.asarray(u, np.float64)
    x = u * (N_DIST_POINTS - 2)
    i = x.astype(np.int64)
    w = x - i
    k = min(max(int(math.floor(log_mr)), 0), N_DIST - 2)
    t = log_mr - k
    lo_row, hi_row = stats[k], stats[k + 1]
    lo = lo_row[i] + w * (lo_row[i + 1] - lo_row[i])
    hi = hi_row[i] + w * (hi_row[i + 1] - hi_row[i])
    llo = np.log10(lo)
    return 10.0 ** (llo + t * (np.log10(hi) - llo))


_MSQ = {}


def mean_sq_jump_steps(log_mr, stats=None):
    """<d^2> of the kernels' draw for one mass ratio, midpoint rule over
    200,000 points (sds.py:172-191), cached per table and mass ratio."""
    st = default_statistics() if stats is None else stats
    key = (id(st), round(float(log_mr), 12))
    if key not in _MSQ:
        u = (np.arange(_MEAN_SQ_POINTS) + 0.5) / _MEAN_SQ_POINTS
        d = icdf_steps_at(st, key[1], u)
        _MSQ[key] = (st, float(np.mean(d * d)))
    return _MSQ[key][1]


# ------------------------------------------------------------ estimators
def _cubic(c, x):
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]))


def diameter_from_mass_nm(mass_amu):
    """Ion diameter (nm) of an ion of this mass at constant density."""
    return MASS_TO_DIAMETER_NM * float(mass_amu) ** (1.0 / 3.0)


def air_K0_from_diameter(d_nm):
    """Reduced mobility in air, cm^2/(V s), from the ion diameter."""
    return 1.0e-5 * 10.0 ** _cubic(_LOGK9_OF_LOGD, math.log10(d_nm))


def diameter_from_air_K0(K0_air):
    """Ion diameter (nm) from a reduced mobility in air."""
    return 10.0 ** _cubic(_LOGD_OF_LOGK9, math.log10(1.0e5 * K0_air))


def gas_over_air(mass_amu, d_nm, gas_mass_amu, gas_d_nm):
    """K0(gas) / K0(air) at a fixed ion size: the Mason-Schamp ratio of
    sqrt(reduced masses) and hard-sphere cross sections pi ((d_ion +
    d_neutral) / 2)^2."""
    m = float(mass_amu)
    mu_air = m * AIR_MASS_AMU / (m + AIR_MASS_AMU)
    mu_gas = m * gas_mass_amu / (m + gas_mass_amu)
    size = (d_nm + AIR_DIAMETER_NM) / (d_nm + gas_d_nm)
    return math.sqrt(mu_air / mu_gas) * size * size


def estimated_K0(mass_amu, gas_mass_amu, gas_d_nm):
    """(K0 in the gas, cm^2/(V s), unit charge; diameter nm) from the
    ion's mass alone."""
    d = diameter_from_mass_nm(mass_amu)
    return air_K0_from_diameter(d) * gas_over_air(mass_amu, d, gas_mass_amu,
                                                  gas_d_nm), d


# ------------------------------------------------- declared species values
SPECIES_KEYS = ("mass_amu", "gas", "charge", "K0_cm2_per_Vs", "d_nm",
                "gas_d_nm")
_SAME_MASS_DA = 1e-6


def _positive_number(v):
    return (isinstance(v, (int, float)) and not isinstance(v, bool)
            and math.isfinite(v) and v > 0)


def species_errors(entries, deck_gas):
    """Refusal messages for collisions.sds_species, each naming its entry
    (sds.py:252-309): a mass and at least one of K0 or the diameter; a K0
    names the gas and the nonzero charge it was measured at; a gas other
    than the deck's needs that gas's diameter for the conversion."""
    from .collision3d import GASES
    if not isinstance(entries, list):
        return [f"collisions.sds_species must be a list of entries, got "
                f"{type(entries).__name__}"]
    errs, masses = [], []
    for i, e in enumerate(entries):
        at = f"collisions.sds_species[{i}]"
        if not isinstance(e, dict):
            errs.append(f"{at} must be an object, got {type(e).__name__}")
            continue
        extra = sorted(set(e) - set(SPECIES_KEYS))
        if extra:
            errs.append(f"{at}: unknown key(s) {extra}; an entry takes "
                        f"{list(SPECIES_KEYS)}")
        m = e.get("mass_amu")
        if not _positive_number(m):
            errs.append(f"{at}: mass_amu = {m!r} must be a finite number > 0")
        elif any(abs(m - s) <= _SAME_MASS_DA for s in masses):
            errs.append(f"{at}: mass {m:g} amu is declared twice")
        else:
            masses.append(m)
        k0, d = e.get("K0_cm2_per_Vs"), e.get("d_nm")
        if k0 is None and d is None:
            errs.append(f"{at}: states neither K0_cm2_per_Vs nor d_nm; give "
                        f"at least one (or remove the entry to estimate the "
                        f"ion from its mass)")
        for key in ("K0_cm2_per_Vs", "d_nm", "gas_d_nm"):
            if e.get(key) is not None and not _positive_number(e[key]):
                errs.append(f"{at}: {key} = {e[key]!r} must be a finite "
                            f"number > 0")
        if k0 is None:
            continue
        z = e.get("charge")
        if not isinstance(z, int) or isinstance(z, bool) or z == 0:
            errs.append(f"{at}: charge = {z!r}. A measured K0 belongs to the "
                        f"charge state it was measured at: state it as a "
                        f"nonzero integer 'charge' (an ion of that charge "
                        f"flies the value as given).")
        gas = e.get("gas")
        if gas is None:
            errs.append(f"{at}: a K0 is specific to the gas it was measured "
                        f"in; state that gas as 'gas'")
        elif gas not in GASES:
            errs.append(f"{at}: gas {gas!r} is not one ion_gym knows "
                        f"({sorted(GASES)})")
        elif gas != deck_gas and e.get("gas_d_nm") is None:
            errs.append(f"{at}: K0 was measured in {gas!r} but the deck's gas "
                        f"is {deck_gas!r}. Converting it needs the diameter "
                        f"of {gas!r}: state it as 'gas_d_nm', or give a K0 "
                        f"measured in {deck_gas!r}.")
    return errs


def species_entry(entries, mass_amu):
    """The sds_species entry declared for this ion mass, or None."""
    m = float(mass_amu)
    for e in entries or ():
        if abs(float(e["mass_amu"]) - m) <= _SAME_MASS_DA:
            return e
    return None


def resolve_K0(mass_amu, gas_name, gas_mass_amu, gas_d_nm, entry):
    """(K0 in the deck's gas per unit charge, cm^2/(V s); diameter nm; how
    it was found) for one ion, in the paper's order (sds.py:312-356): no
    entry, from the mass; a diameter, the air fit at it; a K0, divided by
    the |charge| it was measured at and, when measured in another gas,
    converted at the ion's diameter (found from the K0 by the air fit when
    not declared)."""
    from .collision3d import GASES
    m = float(mass_amu)
    if entry is None:
        K0, d = estimated_K0(m, gas_mass_amu, gas_d_nm)
        return K0, d, "estimated from mass"
    k0_decl, d_decl = entry.get("K0_cm2_per_Vs"), entry.get("d_nm")
    if k0_decl is None:
        d = float(d_decl)
        K0 = air_K0_from_diameter(d) * gas_over_air(m, d, gas_mass_amu,
                                                    gas_d_nm)
        return K0, d, f"from the declared diameter {d:g} nm"
    z_meas = int(entry["charge"])
    k_unit = float(k0_decl) / abs(z_meas)
    meas_gas = entry["gas"]
    same_gas = meas_gas == gas_name
    meas_mass = gas_mass_amu if same_gas else GASES[meas_gas]
    meas_d = gas_d_nm if same_gas else float(entry["gas_d_nm"])
    if d_decl is None:
        ratio0 = gas_over_air(m, diameter_from_mass_nm(m), meas_mass, meas_d)
        d = diameter_from_air_K0(k_unit / ratio0)
    else:
        d = float(d_decl)
    where = f"{meas_gas} at charge {z_meas:+d}"
    if same_gas:
        return k_unit, d, f"declared {float(k0_decl):g} in {where}"
    ratio = (gas_over_air(m, d, gas_mass_amu, gas_d_nm)
             / gas_over_air(m, d, meas_mass, meas_d))
    return (k_unit * ratio, d,
            f"declared {float(k0_decl):g} in {where}, CONVERTED to "
            f"{gas_name} (x {ratio:.4g})")


def declared_K0(mass_amu, gas, K0_cm2_per_Vs, charge=1):
    """A sds_species entry stating a measured K0 (convenience)."""
    return {"mass_amu": float(mass_amu), "gas": gas, "charge": int(charge),
            "K0_cm2_per_Vs": float(K0_cm2_per_Vs)}


# ------------------------------------------------------- per-ion numbers
def transport_params(mass_amu, charge, gas_mass_amu, gas_d_nm, T_k, P_torr,
                     *, gas_name=None, species=None, stats=None):
    """The per-ion SDS numbers at T and P (sds.py:360-407): damping (1/us),
    mfp_mm (the jump scale giving a mean squared jump of 6 D dt per step),
    V_mm_us (mean thermal speed), log_mr_ratio (the table row), K0 and
    its source, K_mm2_per_V_us (|z| K1), D_mm2_per_us (K1 kT/e) and the
    ion diameter.  ``species`` is this ion's sds_species entry or None."""
    m = float(mass_amu)
    if not m > 0.0:
        raise ValueError(f"SDS: ion mass must be positive, got {mass_amu!r}")
    if not (P_torr > 0.0 and T_k > 0.0):
        raise ValueError(f"SDS: T and P must be positive, got T = {T_k!r} K, "
                         f"P = {P_torr!r} Torr")
    if species is not None and gas_name is None:
        raise ValueError("SDS: a declared species value needs the deck's "
                         "gas name to compare its measurement gas with")
    K0, d, how = resolve_K0(m, gas_name, gas_mass_amu, gas_d_nm, species)
    K1 = K0 * 1.0e-4 * (ATM_TORR / P_torr) * (T_k / T_STP_K)   # m^2/(V s)
    m_kg = m * KG_AMU
    D = K1 * K_B * T_k / E_CHG          # m^2/s, numerically mm^2/us
    V = math.sqrt(8.0 * K_B * T_k / (math.pi * m_kg)) * 1.0e-3
    log_mr = math.log10(m / float(gas_mass_amu))
    msq = mean_sq_jump_steps(log_mr, stats)
    return dict(damping=E_CHG / (m_kg * K1) * 1.0e-6,
                mfp_mm=6.0 * D * N_DIST_COLLISIONS / (V * msq),
                V_mm_us=V, log_mr_ratio=log_mr, K0=K0, K0_source=how,
                K_mm2_per_V_us=abs(int(charge)) * K1, D_mm2_per_us=D,
                d_ion_nm=d)


def sds_ion_params(collisions, mass_amu, charge, stats=None):
    """transport_params for one ion on the deck's gas (sds.py:428-447)."""
    from .collision3d import gas_mass
    col = collisions
    if col is None or not getattr(col, "enabled", True):
        raise ValueError("SDS parameters need gas: collisions is None or "
                         "disabled.")
    errs = species_errors(list(col.sds_species or []), col.gas)
    if errs:
        raise ValueError("; ".join(errs))
    return transport_params(
        float(mass_amu), int(charge), gas_mass(col.gas),
        float(col.gas_diam_nm), float(col.T_k), float(col.P_torr),
        gas_name=col.gas, species=species_entry(col.sds_species, mass_amu),
        stats=stats)


def sds_ion_rows(collisions, masses, charge, stats=None):
    """(n, 4) kernel rows (damping, mfp_mm, V_mm_us, log_mr_ratio), one
    transport_params per distinct mass."""
    m = np.asarray(masses, np.float64)
    uniq, inv = np.unique(m, return_inverse=True)
    rows = []
    for v in uniq:
        p = sds_ion_params(collisions, v, charge, stats)
        rows.append((p["damping"], p["mfp_mm"], p["V_mm_us"],
                     p["log_mr_ratio"]))
    return np.asarray(rows, np.float64)[inv]
