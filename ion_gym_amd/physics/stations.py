"""Declared station planes: the kernel planes they become and their
post-hoc hits on recorded trajectories (ion_gym/physics/stations.py).

A station is a plane ``axis = pos_mm`` with optional windows over the two
transverse axes.  Two kinds change a flight and reach the kernels as
planes (fate 5 impact plate: crossings outside the window splat, an empty
window is a wall; fate 6 detector patch, ``on_hit='splat'``: crossings
inside absorb, an empty window is a full-plane detector); a detect
station with ``on_hit='pass'`` is bookkeeping on the records only.  Hits
on records are linear in-step interpolations between bracketing rows.
"""
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

FATE_IMPACT_PLANE = 5
FATE_DETECT = 6
_INF = math.inf
_AXI = {"x": 0, "y": 1, "z": 2}


@dataclass
class StationSpec:
    """ion_gym/io/sim_spec.py:2123-2170: name, kind ('detect' |
    'impact_plane'), on_hit ('pass' | 'splat', detect only), axis, pos_mm
    and windows {axis: [lo, hi]} over the transverse axes."""
    name: str = "station"
    kind: str = "detect"
    on_hit: Optional[str] = None
    axis: str = "x"
    pos_mm: float = 0.0
    window: Dict[str, List[float]] = field(default_factory=dict)


# (kind, on_hit) -> whether the station reaches the kernels
# (stations.py:33-57): impact plates always do, a detector only when it
# absorbs what it sees
_ROUTING = {("impact_plane", None): True, ("detect", "splat"): True,
            ("detect", "pass"): False}


def reaches_kernel(st):
    """True for the stations that become kernel planes, False for pass
    detectors; anything else is refused by name (stations.py:33-57)."""
    kind = getattr(st, "kind", "")
    name = getattr(st, "name", "<unnamed>")
    if kind not in ("impact_plane", "detect"):
        raise ValueError(f"station {name!r}: kind {kind!r} is not a station "
                         f"kind (impact_plane, detect)")
    on_hit = getattr(st, "on_hit", None) if kind == "detect" else None
    routed = _ROUTING.get((kind, on_hit))
    if routed is None:
        raise ValueError(f"station {name!r}: kind='detect' needs on_hit "
                         f"'pass' or 'splat', got {on_hit!r}")
    return routed


def station_planes(spec, world_off_mm=(0.0, 0.0, 0.0)):
    """The spec's kernel-reaching stations as kernel planes (axis index,
    value, sign 0 = both directions, window over the transverse axes in
    xyz order, fate), in the kernel frame field = authored - offset
    (stations.py:59-130).  A window naming the station's own axis, or a
    pair that is not lo <= hi, is refused."""
    off = dict(zip("xyz", [float(v) for v in world_off_mm]))
    out = []
    for st in (getattr(spec, "stations", None) or []):
        if not reaches_kernel(st):
            continue
        sax = st.axis
        t1, t2 = [a for a in "xyz" if a != sax]
        wd = st.window or {}
        if sax in wd:
            raise ValueError(f"station {st.name!r}: window names its own "
                             f"axis {sax!r}; a window constrains the two "
                             f"transverse axes ({t1!r}, {t2!r})")
        for a, w in wd.items():
            if len(w) != 2 or float(w[0]) > float(w[1]):
                raise ValueError(f"station {st.name!r}: window[{a!r}] = "
                                 f"{list(w)!r} is not a (lo, hi) pair")

        def pair(a):
            w = wd.get(a)
            if w is None:
                return (-_INF, _INF)
            return (float(w[0]) - off[a], float(w[1]) - off[a])

        if wd:
            win = (*pair(t1), *pair(t2))
        elif st.kind == "impact_plane":
            win = (_INF, -_INF, _INF, -_INF)
        else:
            win = (-_INF, _INF, -_INF, _INF)
        out.append((_AXI[sax], float(st.pos_mm) - off[sax], 0.0, win,
                    FATE_IMPACT_PLANE if st.kind == "impact_plane"
                    else FATE_DETECT))
    return out


def all_planes(spec, world_off_mm=(0.0, 0.0, 0.0)):
    """The kernel planes of a spec: its declared stations, then any raw
    kernel planes it carries (``spec.planes``)."""
    return station_planes(spec, world_off_mm) + [tuple(p) for p in
                                                 (spec.planes or [])]


def _in_window(window, coords):
    """Boolean mask: every windowed transverse axis of ``coords`` (axis
    name -> value array) inside its [lo, hi]; all True for no window."""
    n = len(next(iter(coords.values())))
    ok = np.ones(n, bool)
    for ax, (lo, hi) in (window or {}).items():
        v = np.asarray(coords[ax], float)
        ok &= (v >= float(lo)) & (v <= float(hi))
    return ok


def station_terminates(st, point):
    """The fate a crossing of station ``st`` at ``point`` (x, y, z)
    assigns, or None when the ion passes (stations.py:133-151): an impact
    plate stops what lands outside its window (everything, with no
    window), an absorbing detector what lands inside it (everything, with
    no window)."""
    if not reaches_kernel(st):
        return None
    wd = st.window or {}
    hit = bool(_in_window(wd, {a: [float(v)] for a, v in
                               zip("xyz", point)})[0])
    if st.kind == "impact_plane":
        return FATE_IMPACT_PLANE if not (wd and hit) else None
    return FATE_DETECT if hit else None


def kernel_stations(spec):
    """The stations that change a flight (stations.py:154-167)."""
    return [st for st in (getattr(spec, "stations", None) or [])
            if reaches_kernel(st)]


_STATE = ("t", "x", "y", "z", "vx", "vy", "vz")


def station_hits(traj, cols, station, *, t_min_us=1.0e-3):
    """Every crossing of one station plane by a recorded trajectory, in
    time order (stations.py:238-275).  Rows j, j+1 bracket a crossing when
    row j lies strictly on one side and row j+1 on the plane or beyond;
    the state (t, x, y, z, vx, vy, vz) at the crossing is the linear
    interpolation between the two rows.  Crossings at t_j <= t_min_us (a
    birth on the plane) are dropped; ``k`` numbers the rest whether or not
    they fall in the window, and ``in_window`` applies the station's
    window."""
    traj = np.asarray(traj, float)
    ix = {c: i for i, c in enumerate(cols)}
    side = traj[:, ix[station.axis]] - float(station.pos_mm)
    lo, hi = side[:-1], side[1:]
    brackets = ((lo < 0) & (hi >= 0)) | ((lo > 0) & (hi <= 0))
    j = np.flatnonzero(brackets)
    j = j[traj[j, ix["t"]] > t_min_us]
    if j.size == 0:
        return []
    frac = (lo[j] / (lo[j] - hi[j]))[:, None]
    sel = [ix[c] for c in _STATE]
    at = traj[j][:, sel] + frac * (traj[j + 1][:, sel] - traj[j][:, sel])
    inside = _in_window(station.window,
                        {c: at[:, n] for n, c in enumerate(_STATE)
                         if c in "xyz"})
    names = ("t_us",) + _STATE[1:]
    return [dict(zip(names, map(float, row)), k=k, in_window=bool(w))
            for k, (row, w) in enumerate(zip(at, inside))]


def _matching_station(spec, name, kind, who):
    """The one declared station with this name / kind (either may be
    None), refusing none or several by name."""
    found = [st for st in (getattr(spec, "stations", None) or [])
             if (kind is None or st.kind == kind)
             and (name is None or st.name == name)]
    if len(found) == 1:
        return found[0]
    what = f"'{kind}' " if kind is not None else ""
    if not found:
        raise ValueError(f"{who}: the spec declares no matching {what}"
                         f"station (name={name!r})")
    raise ValueError(f"{who}: {len(found)} {what}stations match "
                     f"name={name!r}; pass the station name: "
                     f"{[st.name for st in found]}")


def first_station_crossing(traj, cols, spec, *, name=None, kind=None,
                           t_min_us=1.0e-3):
    """The first in-window crossing of the one declared station matching
    ``name`` / ``kind``, or None (stations.py:278-305)."""
    st = _matching_station(spec, name, kind, "first_station_crossing")
    return next((h for h in station_hits(traj, cols, st, t_min_us=t_min_us)
                 if h["in_window"]), None)


def first_detection(traj, cols, spec, *, name=None, t_min_us=1.0e-3):
    """First in-window crossing of the named (or only) detect station
    (stations.py:308-327)."""
    st = _matching_station(spec, name, "detect", "first_detection")
    return next((h for h in station_hits(traj, cols, st, t_min_us=t_min_us)
                 if h["in_window"]), None)
