"""Exact ballistic drift beyond the solved domain, after a grid escape.

Restates ion_gym/physics/drift_extension.py (extend_ballistic, :119-300;
edge_field_max, :88-117) and the r-z hook of
ion_gym/physics/build_rz.py:650-776 for this package's flight spec, whose
stations are the kernel planes (axis, value, sign, window, kind).

When an ion leaves the grid (fate 1) through a face beyond which the deck
declares something -- an enabled bound on that face's axis, or a station
plane on it downstream -- its trajectory gets one appended record at the
exact straight-line crossing of the nearest terminating target (a station
plane that stops the ion, then an enabled bound), or, with only passing
planes ahead, at the furthest reachable plane crossing.  The fate becomes
the station's (5 / 6) or 3 for a bound, and stays 1 otherwise.  The drift
is refused, naming the reason, when the deck has buffer gas (field-free is
not collision-free) or the escape face carries a residual field above
EDGE_FIELD_MAX_V_PER_MM, and nothing is done when no declared target is
reachable within the remaining flight time.
"""
import math

import numpy as np

from .stations import all_planes

EDGE_FIELD_MAX_V_PER_MM = 1e-3

_BASE = ("t", "x", "y", "z", "vx", "vy", "vz")

# field-free extension rule per recorded channel: "const" (unchanged by a
# zero-field drift), "zero" (the drift declares zero field), or
# (previous value, state, dt_us) -> value
_CHANNEL_RULES = {
    "speed": "const", "ke_ev": "const",
    "ke_x": "const", "ke_y": "const", "ke_z": "const",
    "n_col": "const", "wrap_passes": "const",
    "e_field": "zero", "e_axial": "zero", "e_radial": "zero",
    "e_x": "zero", "e_y": "zero", "e_z": "zero",
    "e_axial_tint": "const",
    "path_mm": lambda prev, s, dt: prev + s["sp"] * dt,
    "ke_tint": lambda prev, s, dt: prev + s["ke_ev"] * dt,
}


def edge_field_max(components, axis, side, inset=0):
    """Worst-case |E| over one boundary face: each component a list of
    (array, weight) pairs bounded by sum |w a|, the components combined
    as a vector norm, the max over the face's nodes (``inset`` node rows
    inside the face)."""
    acc = None
    for comp in components:
        mag = None
        for arr, w in comp:
            if arr is None:
                continue
            a = np.asarray(arr, float)
            sl = [slice(None)] * a.ndim
            sl[axis] = int(inset) if side == "lo" else -1 - int(inset)
            face = np.abs(float(w) * a[tuple(sl)])
            mag = face if mag is None else mag + face
        if mag is None:
            continue
        acc = mag ** 2 if acc is None else acc + mag ** 2
    if acc is None:
        raise ValueError("edge_field_max: no field components supplied")
    return float(np.sqrt(acc).max())


def rz_edge_fields(model):
    """Residual |E| (V/mm) on the r-z model's open faces, every drive
    channel at its unit-waveform peak and the gate at full gain (the
    V/m basis converted, build_rz.py:657-669)."""
    gz = getattr(model, "EzG", None)
    gu = getattr(model, "EuG", None)
    cz = ([(model.EzA, 1.0)] + [(model.EzK[k], 1.0)
                                for k in range(model.EzK.shape[0])]
          + [(gz, 1.0)])
    cu = ([(model.EuA, 1.0)] + [(model.EuK[k], 1.0)
                                for k in range(model.EuK.shape[0])]
          + [(gu, 1.0)])
    return {"x_lo": 1e-3 * edge_field_max([cz, cu], 0, "lo"),
            "x_hi": 1e-3 * edge_field_max([cz, cu], 0, "hi"),
            "r_hi": 1e-3 * edge_field_max([cz, cu], 1, "hi")}


def rz_escape_face(x_end, mm_per_gu, nx):
    """(face key, axis, side, words) of an r-z escape at axial x_end: the
    axial ends, else the radial edge, which the r-z bounds convention maps
    to axis 1 (build_rz.py:753-765)."""
    if x_end < 0.0:
        return "x_lo", 0, "lo", "x low (axial entrance)"
    if x_end / mm_per_gu > nx - 1:
        return "x_hi", 0, "hi", "x high (axial exit)"
    return "r_hi", 1, "hi", "radial edge"


def plane_terminates(plane, point):
    """The fate a crossing of station plane (axis, value, sign, window,
    kind) at ``point`` assigns, with the kernels' window rule: an impact
    plate (5) stops the ion outside its window, a detector (6) inside it;
    None when the ion passes."""
    col, _val, _sgn, w, kind = plane
    t1 = point[1] if col == 0 else point[0]
    t2 = point[1] if col == 2 else point[2]
    inside = w[0] <= t1 <= w[1] and w[2] <= t2 <= w[3]
    if int(kind) == 6:
        return 6 if inside else None
    return None if inside else int(kind)


def extend_ballistic(traj, col_names, kind, spec, *, edge_e_vpermm,
                     face_axis, face_side, edge_face="", derived=None,
                     label=""):
    """(traj, kind, info): the trajectory with at most one appended exact
    record (info None when nothing was done, else {dt_us, bnd_face,
    end_mm}).  The input array is not modified."""
    if kind != 1:
        return traj, kind, None
    if face_axis not in (0, 1, 2) or face_side not in ("lo", "hi"):
        raise ValueError(f"drift extension ({label}): face_axis/face_side "
                         f"must name the escape face; got {face_axis!r}/"
                         f"{face_side!r}")
    idx = {c: i for i, c in enumerate(col_names)}
    for c in _BASE:
        if c not in idx:
            raise ValueError(f"drift extension ({label}): trajectory lacks "
                             f"base channel {c!r}")
    last = traj[-1]
    t0 = float(last[idx["t"]])
    pos = np.array([last[idx["x"]], last[idx["y"]], last[idx["z"]]], float)
    vel = np.array([last[idx["vx"]], last[idx["vy"]], last[idx["vz"]]],
                   float)
    budget = float(spec.integration.t_max_us) - (t0 - float(traj[0, idx["t"]]))
    if budget <= 1e-12:
        return traj, kind, None
    bflags, bvals = spec.bounds.as_tuple()
    own_bound = bool(bflags[2 * face_axis + (1 if face_side == "hi" else 0)])
    own_station = any(
        int(p[0]) == face_axis
        and ((face_side == "hi" and float(p[1]) >= pos[face_axis])
             or (face_side == "lo" and float(p[1]) <= pos[face_axis]))
        for p in all_planes(spec))
    if not own_bound and not own_station:
        return traj, kind, None
    tb, bface = math.inf, None
    for k in range(6):
        if not bflags[k]:
            continue
        a, hi = k // 2, (k % 2 == 1)
        v, p, val = vel[a], pos[a], float(bvals[k])
        if v == 0.0:
            continue
        dt = (val - p) / v
        if dt <= 1e-12:
            continue
        crosses = (hi and v > 0.0) or ((not hi) and v < 0.0)
        if crosses and dt < tb:
            tb, bface = dt, k
    st_dts = []
    ts, ts_kind = math.inf, None
    for pl in all_planes(spec):
        a = int(pl[0])
        if vel[a] == 0.0:
            continue
        dt = (float(pl[1]) - pos[a]) / vel[a]
        if dt > 1e-12:
            st_dts.append(dt)
            fate = plane_terminates(pl, pos + vel * dt)
            if fate is not None and dt < ts:
                ts, ts_kind = dt, fate
    bound_reachable = tb <= budget
    term_reachable = ts <= min(tb, budget)
    st_reachable = [d for d in st_dts if d <= min(tb, budget)]
    if not bound_reachable and not st_reachable:
        return traj, kind, None
    col = getattr(spec, "collisions", None)
    if col is not None and getattr(col, "enabled", False):
        raise ValueError(
            f"drift extension ({label}): the deck declares background gas, "
            f"and field-free is not collision-free; cover the drift with "
            f"solved domain, or turn the gas off if it is truly vacuum "
            f"downstream")
    if float(edge_e_vpermm) > EDGE_FIELD_MAX_V_PER_MM:
        raise ValueError(
            f"drift extension ({label}): the escape face "
            f"{edge_face} carries a residual field of "
            f"{float(edge_e_vpermm):.3g} V/mm, above the field-free ceiling "
            f"{EDGE_FIELD_MAX_V_PER_MM:g} V/mm; pad the domain until the "
            f"fringe decays")
    if term_reachable:
        dt_end, kind_out, bface = ts, ts_kind, None
    elif bound_reachable:
        dt_end, kind_out = tb, 3
    else:
        dt_end, kind_out = max(st_reachable), kind
    end = pos + vel * dt_end
    state = {"sp": float(np.linalg.norm(vel)),
             "ke_ev": float(last[idx["ke_ev"]]) if "ke_ev" in idx else None}
    row = np.array(last, float, copy=True)
    row[idx["t"]] = t0 + dt_end
    row[idx["x"]], row[idx["y"]], row[idx["z"]] = end
    derived = derived or {}
    for c in col_names[7:]:
        if c in derived:
            row[idx[c]] = float(derived[c](float(last[idx[c]]), row, idx,
                                           dt_end))
            continue
        rule = _CHANNEL_RULES.get(c)
        if rule == "const":
            continue
        if rule == "zero":
            row[idx[c]] = 0.0
            continue
        if callable(rule):
            if c == "ke_tint" and state["ke_ev"] is None:
                raise ValueError("drift extension: ke_tint is recorded "
                                 "without ke_ev; the extension cannot "
                                 "derive the integrand")
            row[idx[c]] = rule(float(last[idx[c]]), state, dt_end)
            continue
        raise ValueError(f"drift extension ({label}): recorded channel "
                         f"{c!r} has no field-free extension rule")
    return np.vstack([traj, row[None, :]]), kind_out, {
        "dt_us": float(dt_end), "bnd_face": bface,
        "end_mm": tuple(float(v) for v in end)}


def rz_extend(traj, col_names, kind, spec, model, edge_fields, label=""):
    """The r-z hook (build_rz.py:746-770): classify the escape face, then
    extend with the route's radius rule."""
    if kind != 1:
        return traj, kind, None
    fk, fa, fs, words = rz_escape_face(float(traj[-1, 1]), model.mm_per_gu,
                                       model.EzA.shape[0])
    return extend_ballistic(
        traj, col_names, kind, spec, edge_e_vpermm=edge_fields[fk],
        face_axis=fa, face_side=fs, edge_face=words,
        derived={"radius": lambda prev, row, idx, dt:
                 math.hypot(row[idx["y"]], row[idx["z"]])},
        label=label)


def _w3d_max(pack, k):
    """Worst-case |w| of 3-D drive channel k (drift_extension.py:298-311):
    |amp| max|base| + |off|, max|base| 1 for sin/cos/square and the
    table slice's max |value| for a table."""
    kind = int(pack["ch_kind"][k])
    if kind in (3, 4):
        o0, o1 = int(pack["tab_off"][k]), int(pack["tab_off"][k + 1])
        tv = np.asarray(pack["tab_v"], float)[o0:o1]
        base = float(np.max(np.abs(tv))) if tv.size else 0.0
    else:
        base = 1.0
    return (abs(float(pack["ch_amp"][k])) * base
            + abs(float(pack["ch_off"][k])))


def xyz_edge_fields(pack):
    """{face: worst-case residual |E| (V/mm)} on the six faces of a 3-D
    pack (drift_extension.py:313-330, route 3d)."""
    comps = []
    for a, ka in (("EAx", "ExK"), ("EAy", "EyK"), ("EAz", "EzK")):
        K = np.asarray(pack[ka])
        comps.append([(pack[a], 1.0)]
                     + [(K[k], _w3d_max(pack, k)) for k in range(K.shape[0])])
    return {f"{ax}_{s}": edge_field_max(comps, i, s)
            for i, ax in enumerate("xyz") for s in ("lo", "hi")}


def xyz_escape_face(shape, h_mm, p):
    """(face key, axis, side, words) of a 3-D escape at kernel point p:
    the face the point is furthest outside (drift_extension.py:356-400)."""
    best, out = 0.0, None
    for i, ax in enumerate("xyz"):
        hi = (shape[i] - 1) * h_mm
        for side, d in (("lo", -p[i]), ("hi", p[i] - hi)):
            if d > best:
                best, out = d, (f"{ax}_{side}", i, side, f"{ax} {side}")
    if out is None:
        raise ValueError(f"xyz_escape_face: point {tuple(p)} is inside the "
                         f"domain; no escape face")
    return out
