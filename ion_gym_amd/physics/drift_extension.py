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

# field-free extension rule per recorded channel (drift_extension.py:73):
# "const" (unchanged by a zero-field drift), "zero" (the drift declares
# zero field) or "integral" (path length and the KE integral grow with the
# drift time)
_CHANNEL_RULES = {
    "speed": "const", "ke_ev": "const",
    "ke_x": "const", "ke_y": "const", "ke_z": "const",
    "n_col": "const", "wrap_passes": "const",
    "e_field": "zero", "e_axial": "zero", "e_radial": "zero",
    "e_x": "zero", "e_y": "zero", "e_z": "zero",
    "e_axial_tint": "const",
    "path_mm": "integral", "ke_tint": "integral",
}


def _face(arr, axis, side, inset):
    """The node layer of ``arr`` on one face (``inset`` layers inside)."""
    a = np.asarray(arr, float)
    k = int(inset) if side == "lo" else a.shape[axis] - 1 - int(inset)
    return np.take(a, k, axis=axis)


def edge_field_max(components, axis, side, inset=0):
    """Worst-case |E| over one boundary face: each component a list of
    (array, weight) pairs bounded by sum |w a| node by node, the
    components combined as a vector norm, the largest over the face's
    nodes (drift_extension.py:88-117)."""
    bounds = []
    for comp in components:
        terms = [np.abs(float(w) * _face(a, axis, side, inset))
                 for a, w in comp if a is not None]
        if terms:
            bounds.append(np.sum(terms, axis=0))
    if not bounds:
        raise ValueError("edge_field_max: no field components supplied")
    return float(np.linalg.norm(np.stack(bounds), axis=0).max())


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
    planes = all_planes(spec)
    face_slot = 2 * face_axis + (1 if face_side == "hi" else 0)
    ahead = (lambda v: v >= pos[face_axis]) if face_side == "hi" else \
        (lambda v: v <= pos[face_axis])
    if not (bflags[face_slot] or any(int(q[0]) == face_axis
                                     and ahead(float(q[1]))
                                     for q in planes)):
        return traj, kind, None
    tb, bface = _first_bound(pos, vel, bflags, bvals)
    ts, ts_kind, st_dts = _station_crossings(pos, vel, planes)
    horizon = min(tb, budget)
    reach = st_dts[st_dts <= horizon]
    if tb > budget and reach.size == 0:
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
    # the nearest stopping target: a terminating station before any bound,
    # else the bound, else the last passing station within reach
    if ts <= horizon:
        dt_end, kind_out, bface = ts, ts_kind, None
    elif tb <= budget:
        dt_end, kind_out = tb, 3
    else:
        dt_end, kind_out = float(reach.max()), kind
    end = pos + vel * dt_end
    row = _drift_row(traj[-1], col_names, idx, t0 + dt_end, end, vel,
                     dt_end, derived or {}, label)
    return np.vstack([traj, row[None, :]]), kind_out, {
        "dt_us": float(dt_end), "bnd_face": bface,
        "end_mm": tuple(float(v) for v in end)}


_BOUND_AXIS = np.array([0, 0, 1, 1, 2, 2])
_BOUND_HI = np.array([False, True, False, True, False, True])


def _first_bound(pos, vel, flags, vals):
    """(time, bound index) of the first enabled bound the straight line
    from pos with vel crosses outward, or (inf, None)."""
    v = vel[_BOUND_AXIS]
    with np.errstate(divide="ignore", invalid="ignore"):
        dt = (np.asarray(vals, float) - pos[_BOUND_AXIS]) / v
    outward = np.where(_BOUND_HI, v > 0.0, v < 0.0)
    ok = np.asarray(flags, bool) & (v != 0.0) & (dt > 1e-12) & outward
    if not ok.any():
        return math.inf, None
    k = int(np.flatnonzero(ok)[np.argmin(dt[ok])])
    return float(dt[k]), k


def _station_crossings(pos, vel, planes):
    """(time, fate) of the first station crossing that stops the ion (or
    (inf, None)) and the times of every forward crossing."""
    times, first, fate = [], math.inf, None
    for pl in planes:
        a = int(pl[0])
        if vel[a] == 0.0:
            continue
        dt = (float(pl[1]) - pos[a]) / vel[a]
        if dt <= 1e-12:
            continue
        times.append(dt)
        stop = plane_terminates(pl, pos + vel * dt)
        if stop is not None and dt < first:
            first, fate = dt, stop
    return first, fate, np.asarray(times, float)


def _drift_row(last, col_names, idx, t_end, end, vel, dt, derived, label):
    """The appended record: the base channels at the end of the straight
    drift, each optional channel by its field-free rule (a caller's
    ``derived`` rule first)."""
    speed = float(np.linalg.norm(vel))
    ke = float(last[idx["ke_ev"]]) if "ke_ev" in idx else None
    base = dict(zip(_BASE, (t_end, *end, *vel)))
    out = np.empty(len(col_names))
    for j, c in enumerate(col_names):
        if c in base:
            out[j] = base[c]
            continue
        prev = float(last[j])
        rule = _CHANNEL_RULES.get(c)
        if c in derived:
            row_so_far = np.array(last, float, copy=True)
            for b, v in base.items():
                row_so_far[idx[b]] = v
            out[j] = float(derived[c](prev, row_so_far, idx, dt))
        elif rule == "const":
            out[j] = prev
        elif rule == "zero":
            out[j] = 0.0
        elif c == "path_mm":
            out[j] = prev + speed * dt
        elif c == "ke_tint":
            if ke is None:
                raise ValueError("drift extension: ke_tint is recorded "
                                 "without ke_ev; the extension cannot "
                                 "derive the integrand")
            out[j] = prev + ke * dt
        else:
            raise ValueError(f"drift extension ({label}): recorded channel "
                             f"{c!r} has no field-free extension rule")
    return out


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
