"""Per-ion hard-sphere cross sections.

The flight kernels take one cross section per ion (``sigma_m2`` in the ion
arrays).  ``ion_sigma_m2(collisions, mz, charge)`` is where it comes from,
as in ion_gym/physics/ccs.py:157-192, selected by
``collisions.sigma_model``:

``"fixed"``
    ``collisions.sigma_m2`` for every ion.
``"ccs_fit"`` (the default, ccs.py:67)
    the power law ``CCS = A (m/z)^B`` (A^2) fitted to drift-tube cross
    sections in N2 (Picache et al., Chem. Sci. 2019, 10, 983-993: the
    Unified CCS Compendium, singly charged ions), one row per fit class
    (``collisions.ccs_fit_class``, "all" by default), evaluated at the
    ion's mass for any charge.  The coefficient table
    ``ccs_fit_n2_z1.json`` is ion_gym's, kept with its numbers so that the
    same ion flies the same cross section.

Rules kept from the reference: vacuum (no collisions, or disabled) gives
0.0; a gas other than N2 falls back to ``collisions.sigma_m2`` with an
advisory that names the gas, the value flown and how to state it on
purpose; an m/z outside the fitted range is extrapolated with an advisory;
an unknown model or fit class is refused by name (``CCSModelRefused``, a
ValueError).  Advisories are ``IonGymAdvisory`` warnings, so Python's
warning filter governs repeats (each distinct text once by default).
"""
import json
import os
import warnings

SIGMA_MODELS = ("fixed", "ccs_fit")
SIGMA_MODEL_DEFAULT = "ccs_fit"
# SIMION's collision_hs1.lua value (He with a ~200 amu ion), still the
# reference's CollisionSpec.sigma_m2 default
LEGACY_FIXED_SIGMA_M2 = 2.27e-18
CCS_FIT_GAS = "N2"
A2_TO_M2 = 1.0e-20
_TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                           "ccs_fit_n2_z1.json")


class IonGymAdvisory(UserWarning):
    """A stated approximation or fallback of a flight (not an error)."""


class CCSModelRefused(ValueError):
    """A cross-section request the selected model cannot honour."""


_TABLE = None


def load_fit_table():
    """The coefficient table: its header and ``by_label`` {class: row};
    each row has A_A2, B, n, resid_sd_frac, mz_min and mz_max."""
    global _TABLE
    if _TABLE is None:
        with open(_TABLE_PATH, encoding="utf-8") as fh:
            t = json.load(fh)
        t["by_label"] = {row["label"]: row for row in t["fits"]}
        _TABLE = t
    return _TABLE


def fit_classes():
    """The labels ``collisions.ccs_fit_class`` may name."""
    return tuple(load_fit_table()["by_label"])


def _row(fit_class):
    rows = load_fit_table()["by_label"]
    try:
        return rows[fit_class]
    except KeyError:
        raise CCSModelRefused(
            f"collisions.ccs_fit_class {fit_class!r} is not in the CCS fit "
            f"table. Known classes: {sorted(rows)}") from None


def ccs_fit_A2(mz, fit_class="all"):
    """(CCS in A^2, extrapolated?, fit row) at one m/z."""
    row = _row(fit_class)
    m = float(mz)
    if not m > 0.0:
        raise CCSModelRefused(f"m/z must be positive, got {m}")
    inside = row["mz_min"] <= m <= row["mz_max"]
    return row["A_A2"] * m ** row["B"], not inside, row


def _enabled(col):
    return col is not None and bool(getattr(col, "enabled", True))


def hs_sigma_in_play(col):
    """Whether a flight reads a hard-sphere cross section: gas on and a
    model other than SDS (SDS never reads one)."""
    return _enabled(col) and str(getattr(col, "model", "hs")).lower() != "sds"


def fit_domain_reason(col, charge=1):
    """Why the fit does not apply to this deck's gas, or None.  Charge is
    no limit: the fit is taken at the ion's mass (see charge_note)."""
    gas = str(col.gas)
    if gas == CCS_FIT_GAS:
        return None
    return (f"the CCS fit is drift-tube data in {CCS_FIT_GAS}, and the "
            f"deck's gas is {gas!r}")


def fallback_text(reason, sigma_m2):
    """The wording of the static-sigma fallback."""
    s = float(sigma_m2)
    legacy = (" This is the legacy default (SIMION's He value for a ~200 amu "
              "ion), not a value the deck chose."
              if s == LEGACY_FIXED_SIGMA_M2 else "")
    return (f"collisions.sigma_model 'ccs_fit' does not apply: {reason}. "
            f"The hard-sphere flight falls back to the static cross section "
            f"collisions.sigma_m2 = {s:.4g} m^2 for every ion.{legacy} State "
            f"sigma_model 'fixed' (with a sigma_m2 for this gas and charge) "
            f"to fly a static cross section on purpose.")


def _advise(text):
    warnings.warn(text, IonGymAdvisory, stacklevel=3)


def ion_sigma_m2(collisions, mz, charge):
    """The hard-sphere cross section (m^2) of one ion of mass ``mz`` (Da)
    and signed ``charge`` under ``collisions`` (None or disabled: 0.0)."""
    col = collisions
    if not _enabled(col):
        return 0.0
    model = str(getattr(col, "sigma_model", SIGMA_MODEL_DEFAULT))
    if model == "fixed":
        return float(col.sigma_m2)
    if model != "ccs_fit":
        raise CCSModelRefused(
            f"collisions.sigma_model {model!r} is unknown. Known models: "
            f"{SIGMA_MODELS}")
    reason = fit_domain_reason(col, charge)
    if reason is not None:
        if hs_sigma_in_play(col):
            _advise(f"[CCS] {fallback_text(reason, col.sigma_m2)}")
        return float(col.sigma_m2)
    fit_class = str(getattr(col, "ccs_fit_class", "all"))
    a2, outside, row = ccs_fit_A2(mz, fit_class)
    if outside:
        _advise(f"[CCS] m/z {float(mz):g} lies outside the fitted range "
                f"{row['mz_min']:.0f} to {row['mz_max']:.0f} of CCS fit class "
                f"{fit_class!r} (n = {row['n']}). The power law is "
                f"EXTRAPOLATED to CCS = {a2:.1f} A^2. Fit scatter inside the "
                f"range is {100 * row['resid_sd_frac']:.1f} %, and nothing "
                f"bounds the error outside it.")
    return a2 * A2_TO_M2


def ion_sigmas_m2(collisions, masses, charge):
    """ion_sigma_m2 over an array of masses, each distinct mass looked up
    once (so an advisory is raised once per mass, not per ion)."""
    import numpy as np
    m = np.asarray(masses, np.float64)
    uniq, inv = np.unique(m, return_inverse=True)
    vals = np.array([ion_sigma_m2(collisions, v, charge) for v in uniq])
    return vals[inv].reshape(m.shape)


def validate_sigma_model(collisions):
    """Refusal messages for a deck whose sigma model cannot mean what it
    asks: an unknown model (always), a fit class not in the table (only
    when a hard-sphere cross section is flown under 'ccs_fit')."""
    model = str(getattr(collisions, "sigma_model", SIGMA_MODEL_DEFAULT))
    if model not in SIGMA_MODELS:
        return [f"collisions.sigma_model {model!r} is unknown. Known models "
                f"are {SIGMA_MODELS}"]
    if model != "ccs_fit" or not hs_sigma_in_play(collisions):
        return []
    try:
        _row(str(getattr(collisions, "ccs_fit_class", "all")))
    except CCSModelRefused as e:
        return [str(e)]
    return []


def _fit_applies(col):
    return (hs_sigma_in_play(col)
            and str(getattr(col, "sigma_model", SIGMA_MODEL_DEFAULT))
            == "ccs_fit")


def fallback_note(spec):
    """'WARNING: ...' when a deck flies the fit's static fallback, else
    None."""
    col = spec.collisions
    if not _fit_applies(col):
        return None
    reason = fit_domain_reason(col, spec.source.charge)
    if reason is None:
        return None
    return "WARNING: " + fallback_text(reason, col.sigma_m2)


def charge_note(spec):
    """The approximation a multiply charged ion flies under the fit, or
    None (|z| = 1, no fit in play, or the fallback)."""
    z = int(spec.source.charge)
    col = spec.collisions
    if abs(z) == 1 or not _fit_applies(col) or fit_domain_reason(col, z):
        return None
    return (f"collisions: the CCS fit is of singly charged ions; the charge "
            f"{z:+d} ions fly the fit at their mass, so the growth of the "
            f"cross section with charge is neglected and their mobility is "
            f"an upper estimate.")


def extrapolation_note(spec):
    """'WARNING: the CCS fit ...' naming the masses flown outside the
    fitted range, or None."""
    col = spec.collisions
    if not _fit_applies(col) or fit_domain_reason(col, spec.source.charge):
        return None
    fc = str(getattr(col, "ccs_fit_class", "all"))
    out = [float(m) for m in spec.source.mz_list
           if ccs_fit_A2(m, fc)[1]]
    if not out:
        return None
    row = _row(fc)
    names = ", ".join(f"m/z {m:g}" for m in out)
    return (f"WARNING: the CCS fit (class {fc!r}, fitted over m/z "
            f"{row['mz_min']:.0f} to {row['mz_max']:.0f}) is EXTRAPOLATED "
            f"for {names}.")
