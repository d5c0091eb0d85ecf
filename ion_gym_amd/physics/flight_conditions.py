"""Voltages, drives, ion, gas and integration set after the solve
(ion_gym/physics/flight_conditions.py:55-243).

The bases are per electrode, so these conditions re-weight cached bases
and never re-solve.  ``set_flight_conditions`` applies each named setting
to the deck, refuses a name the deck does not have, and refuses the whole
call when the deck's basis-cache key (io.basis_cache.key) no longer
matches the solved fields: the key at entry, or the ``solved_key`` the
caller captured right after the solve (which also catches a geometry
change made outside the call).  Returns the report lines, printed when
``show``.
"""
from dataclasses import fields as _fields

from ..io.basis_cache import key as _basis_key

__all__ = ["set_flight_conditions", "solved_key", "FlightConditionError"]


def solved_key(spec):
    """The geometry key of the fields just solved for ``spec``."""
    return _basis_key(spec)


class FlightConditionError(ValueError):
    """A condition names something the deck lacks, or would move the
    geometry out from under the solved fields."""


def _names(obj):
    return sorted(f.name for f in _fields(obj))


def _assign(report, label, obj, attr, value, missing):
    if not hasattr(obj, attr):
        raise FlightConditionError(missing)
    report.append(f"{label}{attr}: {getattr(obj, attr)!r} -> {value!r}")
    setattr(obj, attr, value)


def _set_voltages(spec, voltages, report):
    g = spec.geometry
    by_name = {e.name: e for e in g.electrodes}
    ladders = {d.name: d for d in g.dc_groups}
    for name, v in voltages.items():
        el = by_name.get(name)
        if el is None:
            raise FlightConditionError(
                f"voltage for unknown electrode {name!r}; this deck has "
                f"{sorted(by_name)}")
        v = float(v)
        grp = ladders.get(el.dc_group) if el.dc_group else None
        if grp is None:
            report.append(f"V[{name}]: {el.dc:g} -> {v:g} V")
            el.dc = v
            continue
        mates = [e for e in g.electrodes if e.dc_group == el.dc_group]
        if not (grp.uniform or len(mates) == 1):
            raise FlightConditionError(
                f"voltage for {name!r} cannot be set per-electrode: it is a "
                f"member of the non-uniform dc ladder {el.dc_group!r} "
                f"({len(mates)} members), whose taps are DERIVED from the "
                f"group's v_in/v_out by resolve_dc_groups() at every build. "
                f"Setting one member's dc would be silently overwritten. "
                f"Drive the ladder instead: set_flight_conditions(spec, "
                f"dc_groups={{'{el.dc_group}': {{'v_in': ..., "
                f"'v_out': ...}}}}).")
        others = len(mates) - 1
        report.append(f"V[{name}] via uniform dc_group {el.dc_group!r}: "
                      f"{grp.v_in:g} -> {v:g} V"
                      + (f" (also moves {others} other member(s))"
                         if others else ""))
        grp.v_in = grp.v_out = v
        for e in mates:
            e.dc = v


def _set_group_fields(groups, settings, report, what):
    for name, values in settings.items():
        grp = groups.get(name)
        if grp is None:
            raise FlightConditionError(
                f"{what} settings for unknown group {name!r}; this deck "
                f"has {sorted(groups)}")
        for attr, val in values.items():
            extra = (f"; settable: {_names(grp)}" if what == "dc_group"
                     else "")
            _assign(report, f"{name}.", grp, attr, val,
                    f"{'dc_group' if what == 'dc_group' else 'drive group'} "
                    f"{name!r} has no field {attr!r}{extra}")


def _set_ion(spec, ion, report):
    s = spec.source
    for attr, val in ion.items():
        if attr == "ke_ev":
            lo, hi = float(val[0]), float(val[1])
            report.append(f"ion KE: {s.ke_lo:g}-{s.ke_hi:g} -> "
                          f"{lo:g}-{hi:g} eV")
            s.ke_lo, s.ke_hi = lo, hi
            continue
        _assign(report, "ion ", s, attr, val,
                f"source has no field {attr!r}; settable: {_names(s)}")


def _set_gas(spec, gas, report):
    c = spec.collisions
    if c is None:
        raise FlightConditionError("gas settings on a deck without a "
                                   "collisions block")
    settable = set(_names(c)) - {"P_torr"}
    unknown = sorted(k for k in gas
                     if k not in settable and k not in ("p_torr", "P_torr"))
    if unknown:
        raise FlightConditionError(
            f"gas has no field(s) {unknown}"
            + (" (P_pa is deprecated: give p_torr)" if "P_pa" in unknown
               else "")
            + f"; settable: {['p_torr'] + sorted(settable)}")
    if "p_torr" in gas and "P_torr" in gas:
        raise FlightConditionError("gas gives both p_torr and P_torr; give "
                                   "the pressure once")
    torr = gas.get("p_torr", gas.get("P_torr"))
    if torr is not None:
        report.append(f"gas pressure: {c.P_torr!r} -> {float(torr):g} Torr")
        c.P_torr = float(torr)
    for k, val in gas.items():
        if k not in ("p_torr", "P_torr"):
            _assign(report, "gas ", c, k, val, f"gas has no field {k!r}")


def set_flight_conditions(spec, *, voltages=None, drives=None, ion=None,
                          gas=None, integration=None, dc_groups=None,
                          solved_key=None, show=True):
    """Apply post-solve conditions to ``spec`` and return the report.

    voltages     {electrode name: volts} -- the DC of a named electrode (a
                 member of a uniform DC group moves the group; a member of
                 a graded ladder is refused, naming dc_groups)
    dc_groups    {group: {v_in=, v_out=, interp=, uniform=, mode=}}
    drives       {group: {amplitude_v=, frequency_hz=, phase_deg=,
                          offset_v=, duty=, waveform=, ...}}
    ion          {mz_list=, charge=, ke_ev=(lo, hi), n_ions=, ...} (any
                 SourceSpec field)
    gas          {p_torr= (or P_torr=), T_k=, model=, sigma_model=, ...}
                 (any CollisionSpec field)
    integration  {dt_ns=, t_max_us=, rec_every=, max_records=, ...}

    Raises FlightConditionError on an unknown name or when the geometry
    key no longer matches the solved fields."""
    before = solved_key if solved_key is not None else _basis_key(spec)
    report = []
    if voltages:
        _set_voltages(spec, voltages, report)
    if dc_groups:
        _set_group_fields({d.name: d for d in spec.geometry.dc_groups},
                          dc_groups, report, "dc_group")
    if drives:
        _set_group_fields({d.name: d for d in spec.geometry.rf_groups},
                          drives, report, "drive")
    if ion:
        _set_ion(spec, ion, report)
    if gas:
        _set_gas(spec, gas, report)
    if integration:
        for attr, val in integration.items():
            _assign(report, "", spec.integration, attr, val,
                    f"integration has no field {attr!r}")
    after = _basis_key(spec)
    if after != before:
        raise FlightConditionError(
            "the geometry key does not match the solved fields "
            f"({before} -> {after}), so the fields shown above no longer "
            "describe this spec and the next build_run would RE-SOLVE. "
            "Flight conditions are voltage/ion/gas/integration only. If the "
            "geometry really must change, change it in the geometry cell and "
            "re-solve deliberately."
            + ("" if solved_key is not None else
               "  (No solved_key was passed, so this call could only compare "
               "against the key at entry -- pass solved_key=solved_key(spec) "
               "captured right after build_run to catch changes made "
               "outside this call too.)"))
    if show:
        print("FLIGHT CONDITIONS (applied to the SOLVED model; the next "
              "build_run re-weights cached bases, it does not re-solve)")
        for line in report:
            print("  " + line)
        if not report:
            print("  (none changed -- flying the deck's own conditions)")
    return report
