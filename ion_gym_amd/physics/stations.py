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


def reaches_kernel(st):
    """True for the stations that become kernel planes, False for pass
    detectors; anything else is refused by name (stations.py:33-57)."""
    kind = getattr(st, "kind", "")
    name = getattr(st, "name", "<unnamed>")
    if kind == "impact_plane":
        return True
    if kind == "detect":
        on_hit = getattr(st, "on_hit", None)
        if on_hit == "splat":
            return True
        if on_hit == "pass":
            return False
        raise ValueError(f"station {name!r}: kind='detect' needs on_hit "
                         f"'pass' or 'splat', got {on_hit!r}")
    raise ValueError(f"station {name!r}: kind {kind!r} is not a station kind "
                     f"(impact_plane, detect)")


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


def station_terminates(st, point):
    """The fate a crossing of station ``st`` at ``point`` (x, y, z)
    assigns, or None when the ion passes (stations.py:133-151)."""
    if not reaches_kernel(st):
        return None
    wd = st.window or {}
    p = dict(zip("xyz", (float(v) for v in point)))
    inside = bool(wd) and all(float(w[0]) <= p[a] <= float(w[1])
                              for a, w in wd.items())
    if st.kind == "impact_plane":
        return None if inside else FATE_IMPACT_PLANE
    return FATE_DETECT if (inside or not wd) else None


def kernel_stations(spec):
    """The stations that change a flight (stations.py:154-167)."""
    return [st for st in (getattr(spec, "stations", None) or [])
            if reaches_kernel(st)]


_STATE = ("t", "x", "y", "z", "vx", "vy", "vz")


def station_hits(traj, cols, station, *, t_min_us=1.0e-3):
    """Every crossing of one station plane by a recorded trajectory, in
    time order (stations.py:238-275): a crossing between rows j and j+1
    when row j is strictly on one side and row j+1 on the plane or past
    it; the state is interpolated linearly at the crossing; k counts all
    crossings (in or out of the window); crossings at t <= t_min_us are
    dropped (a birth on the plane); ``in_window`` applies the station's
    windows."""
    traj = np.asarray(traj, float)
    ix = {c: i for i, c in enumerate(cols)}
    t = traj[:, ix["t"]]
    q = traj[:, ix[station.axis]] - float(station.pos_mm)
    a, b = q[:-1], q[1:]
    crossing = ((a < 0) & (b >= 0)) | ((a > 0) & (b <= 0))
    hits = []
    k = 0
    for j in np.nonzero(crossing)[0]:
        if t[j] <= t_min_us:
            continue
        f = q[j] / (q[j] - q[j + 1])
        st = {c: float(traj[j, ix[c]] + f * (traj[j + 1, ix[c]]
                                             - traj[j, ix[c]]))
              for c in _STATE}
        st["t_us"] = st.pop("t")
        st["k"] = k
        k += 1
        inside = True
        for wax, (lo, hi) in (station.window or {}).items():
            if not (float(lo) <= st[wax] <= float(hi)):
                inside = False
                break
        st["in_window"] = inside
        hits.append(st)
    return hits


def first_station_crossing(traj, cols, spec, *, name=None, kind=None,
                           t_min_us=1.0e-3):
    """The first in-window crossing of the one declared station matching
    ``name`` / ``kind``, or None (stations.py:278-305)."""
    sts = [st for st in (spec.stations or [])
           if (kind is None or st.kind == kind)
           and (name is None or st.name == name)]
    want = f"'{kind}' " if kind is not None else ""
    if not sts:
        raise ValueError(f"first_station_crossing: the spec declares no "
                         f"matching {want}station (name={name!r})")
    if len(sts) > 1:
        raise ValueError(f"first_station_crossing: {len(sts)} {want}stations "
                         f"match name={name!r}; pass the station name: "
                         f"{[d.name for d in sts]}")
    for h in station_hits(traj, cols, sts[0], t_min_us=t_min_us):
        if h["in_window"]:
            return h
    return None


def first_detection(traj, cols, spec, *, name=None, t_min_us=1.0e-3):
    """First in-window crossing of the named (or only) detect station
    (stations.py:308-327)."""
    try:
        return first_station_crossing(traj, cols, spec, name=name,
                                      kind="detect", t_min_us=t_min_us)
    except ValueError as e:
        if "declares no matching" in str(e):
            raise ValueError(
                "first_detection: the spec declares no matching 'detect' "
                f"station (name={name!r})") from e
        raise
