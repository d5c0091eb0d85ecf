"""Ensemble statistics of a flight (ion_gym/physics/stats.py:45-447).

``compute_stats(results, transmitted_fate=...)`` reduces an ensemble's
summaries: fate counts and transmission (the success fate is a required
argument, ``auto_transmitted_fate(spec)`` answers it for transmission
devices), per-mass time-of-flight mean, sigma, FWHM and resolving power
of the transmitted ions, end-point spreads, mean collision count, and
where ions landed -- grouped by declared plane when the plane set is
given (``enabled_planes(spec)``), with the plane's normal known, else one
group per fate with the normal taken as the axis of smallest spread.
Standard deviations are population (ddof = 0) throughout, as the
reference states.
"""
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .ion_envelope import FATE_NAME

FWHM_PER_SIGMA = 2.35482004503


def _summaries(results):
    return [r.summary if hasattr(r, "summary") else r for r in results]


def _mean_std(xs):
    xs = [float(x) for x in xs if x is not None]
    n = len(xs)
    if n == 0:
        return None, None, 0
    mu = sum(xs) / n
    if n == 1:
        return mu, 0.0, 1
    var = sum((x - mu) ** 2 for x in xs) / n
    return mu, math.sqrt(var), n


def auto_transmitted_fate(spec):
    """3 when any bounding plane is enabled, else 1 (stats.py:69-89)."""
    b = getattr(spec, "bounds", None)
    if b is not None:
        for ax in ("x", "y", "z"):
            for side in ("min", "max"):
                if getattr(b, f"{ax}_{side}_on", False):
                    return 3
    return 1


def mz_of_results(spec, n):
    """Per-ion mass for n results, contiguous blocks of n_ions
    (stats.py:92-108)."""
    lst = list(spec.source.mz_list or [])
    if not lst:
        return [None] * n
    n_per = max(1, int(getattr(spec.source, "n_ions", 1)))
    return [lst[(i // n_per) % len(lst)] for i in range(n)]


def enabled_planes(spec):
    """The enabled bounding planes as (label, axis, coord_mm)
    (stats.py:186-201)."""
    b = getattr(spec, "bounds", None) if spec is not None else None
    if b is None:
        return []
    out = []
    for axis in ("x", "y", "z"):
        for side in ("min", "max"):
            if getattr(b, f"{axis}_{side}_on", False):
                coord = float(getattr(b, f"{axis}_{side}"))
                out.append((f"{axis}={coord:+g}", axis, coord))
    return out


@dataclass
class MzStats:
    mz: Optional[float]
    n: int
    tof_mean: Optional[float]
    tof_sigma: Optional[float]
    tof_fwhm: Optional[float]
    resolving_power: Optional[float]


@dataclass
class PlaneStats:
    fate: int
    n: int
    normal: str
    normal_mean: float
    normal_sigma: float
    in_plane: tuple
    centroid: tuple
    sigma: tuple
    r_mean: float
    r_rms: float
    r_p95: float
    plane_label: str = ""
    ke_n: int = 0
    ke_mean: float = 0.0
    ke_sigma: float = 0.0
    ke_min: float = 0.0
    ke_max: float = 0.0
    v_norm_mean: float = 0.0
    v_norm_sigma: float = 0.0
    v_perp_mean: float = 0.0
    v_perp_sigma: float = 0.0
    v_perp_rms: float = 0.0
    div_mrad_rms: float = 0.0
    div_mrad_p95: float = 0.0
    tof_n: int = 0
    tof_mean: float = 0.0
    tof_fwhm: float = 0.0
    tof_R: float = 0.0
    tof_note: str = ""


@dataclass
class Stats:
    n_total: int
    fate_counts: dict
    transmitted_fate: int
    transmission: Optional[float]
    per_mz: list
    end_spread: dict
    r_end_mean: Optional[float]
    r_end_sigma: Optional[float]
    n_col_mean: Optional[float]
    impact_spread: dict
    n_transmitted: int
    landing: dict = None


def _tof_block(rows):
    """Arrival-time statistics of one landing group (stats.py:204-223): a
    stated blank for mixed masses or no recorded tof; no width below two
    ions."""
    mzs = {s.get("mz") for s in rows if s.get("mz") is not None}
    tofs = [s.get("tof") for s in rows if s.get("tof") is not None]
    if not tofs:
        return dict(tof_n=0, tof_note="route records no tof")
    if len(mzs) > 1:
        return dict(tof_n=0, tof_note="mixed masses — see the per-mass tabs")
    mu, sd, cnt = _mean_std(tofs)
    if cnt < 2 or not sd:
        return dict(tof_n=cnt, tof_mean=mu, tof_note="n < 2 — no width")
    fwhm = FWHM_PER_SIGMA * sd
    return dict(tof_n=cnt, tof_mean=mu, tof_fwhm=fwhm, tof_R=mu / (2 * fwhm))


def _plane_stats(rows, fate, normal=None, plane_label=""):
    """Landing statistics of one group (stats.py:226-297)."""
    pts = {a: [s.get(f"{a}_end") for s in rows] for a in "xyz"}
    if not rows or any(v is None for v in pts["x"]):
        return None
    stat = {}
    for a in "xyz":
        mu, sd, cnt = _mean_std(pts[a])
        if cnt == 0:
            return None
        stat[a] = (mu, sd or 0.0)
    if normal not in ("x", "y", "z"):
        normal = min("xyz", key=lambda a: stat[a][1])
    ip = tuple(a for a in "xyz" if a != normal)
    cx, cy = stat[ip[0]][0], stat[ip[1]][0]
    rs = sorted(math.hypot(px - cx, py - cy)
                for px, py in zip(pts[ip[0]], pts[ip[1]]))
    n = len(rs)
    r_mean = sum(rs) / n
    r_rms = math.sqrt(sum(r * r for r in rs) / n)
    r_p95 = rs[min(n - 1, int(0.95 * n))]
    kes = [r.get("ke_end") for r in rows if r.get("ke_end") is not None]
    ke_mu, ke_sd, _ = _mean_std(kes) if kes else (0.0, 0.0, 0)
    vel = [(r.get("vx_end"), r.get("vy_end"), r.get("vz_end")) for r in rows]
    vel = [v for v in vel if all(c is not None for c in v)]
    vn_mu = vn_sd = vp_mu = vp_sd = vp_rms = d_rms = d_p95 = 0.0
    if vel:
        V = np.asarray(vel, float)
        k = "xyz".index(normal)
        vN = V[:, k]
        vP = np.hypot(*[V[:, j] for j in range(3) if j != k])
        vn_mu = float(vN.mean())
        vn_sd = float(vN.std() if len(vN) > 1 else 0.0)
        vp_mu = float(vP.mean())
        vp_sd = float(vP.std() if len(vP) > 1 else 0.0)
        vp_rms = float(np.sqrt((vP ** 2).mean()))
        with np.errstate(invalid="ignore", divide="ignore"):
            ang = np.abs(np.arctan2(vP, np.abs(vN))) * 1e3
        ang = ang[np.isfinite(ang)]
        if ang.size:
            d_rms = float(np.sqrt((ang ** 2).mean()))
            d_p95 = float(np.percentile(ang, 95.0))
    return PlaneStats(
        fate=fate, n=n, normal=normal, normal_mean=stat[normal][0],
        normal_sigma=stat[normal][1], in_plane=ip, centroid=(cx, cy),
        sigma=(stat[ip[0]][1], stat[ip[1]][1]), r_mean=r_mean, r_rms=r_rms,
        r_p95=r_p95, plane_label=plane_label, ke_n=len(kes),
        ke_mean=ke_mu or 0.0, ke_sigma=ke_sd or 0.0,
        ke_min=min(kes) if kes else 0.0, ke_max=max(kes) if kes else 0.0,
        v_norm_mean=vn_mu, v_norm_sigma=vn_sd, v_perp_mean=vp_mu,
        v_perp_sigma=vp_sd, v_perp_rms=vp_rms, div_mrad_rms=d_rms,
        div_mrad_p95=d_p95, **_tof_block(rows))


def compute_stats(results, *, transmitted_fate=None, mz=None, planes=None,
                  plane_tol_mm=2.0):
    """Reduce an ensemble (driver results or summary dicts) to Stats
    (stats.py:315-447).  transmitted_fate is required; mz optional per
    result; planes the declared (label, axis, coord) set, whose match
    tolerance is capped at half the smallest gap between planes on one
    axis."""
    if transmitted_fate is None:
        raise ValueError(
            "compute_stats: transmitted_fate is required -- state which fate "
            "counts as success: auto_transmitted_fate(spec) for a "
            "transmission device (3 = bounding plane, 1 = left the domain), "
            "or 2 only when being held to t_max is success (a trap).")
    if transmitted_fate not in FATE_NAME:
        raise ValueError(f"compute_stats: transmitted_fate "
                         f"{transmitted_fate!r} is not a fate code")
    S = _summaries(results)
    n = len(S)
    mz = list(mz) if mz is not None else [None] * n
    fate_counts = {}
    for s in S:
        k = s.get("kind")
        fate_counts[k] = fate_counts.get(k, 0) + 1
    ok = [(s, m) for s, m in zip(S, mz) if s.get("kind") == transmitted_fate]
    impact_rows = [s for s in S if s.get("kind") == 0]
    transmission = (len(ok) / n) if n else None
    groups = {}
    for s, m in ok:
        groups.setdefault(m, []).append(s.get("tof"))
    per_mz = []
    for m in sorted(groups, key=lambda v: (v is None, v)):
        mu, sd, cnt = _mean_std(groups[m])
        fwhm = R = None
        if mu is not None and sd is not None:
            fwhm = FWHM_PER_SIGMA * sd
            R = (mu / (2 * fwhm)) if fwhm > 0 else float("inf")
        per_mz.append(MzStats(m, cnt, mu, sd, fwhm, R))

    def spread(rows):
        out = {}
        for key in ("x_end", "y_end", "z_end"):
            mu, sd, cnt = _mean_std([s.get(key) for s in rows])
            if cnt:
                out[key[0]] = (mu, sd)
        return out

    ok_rows = [s for s, _ in ok]
    r_mu, r_sd, _ = _mean_std([s.get("r_end") for s in ok_rows])
    nc_mu, _, _ = _mean_std([s.get("n_col") for s in S])
    landing = {}
    plist = list(planes) if planes else []
    tol = float(plane_tol_mm)
    for ax in "xyz":
        cs = sorted(c for _, a, c in plist if a == ax)
        gaps = [b - a for a, b in zip(cs, cs[1:]) if b > a]
        if gaps:
            tol = min(tol, 0.5 * min(gaps))
    for code in sorted(fate_counts, key=lambda c: (c is None, c)):
        rows = [s for s in S if s.get("kind") == code]
        if not plist:
            ps = _plane_stats(rows, code)
            if ps is not None:
                landing[code] = ps
            continue
        by_plane, residual = {}, []
        for s in rows:
            best = None
            for label, axis, coord in plist:
                v = s.get(f"{axis}_end")
                if v is None:
                    continue
                d = abs(v - coord)
                if best is None or d < best[0]:
                    best = (d, label, axis)
            if best is not None and best[0] <= tol:
                by_plane.setdefault((best[1], best[2]), []).append(s)
            else:
                residual.append(s)
        for (label, axis), prows in sorted(by_plane.items()):
            ps = _plane_stats(prows, code, normal=axis, plane_label=label)
            if ps is not None:
                landing[f"{code}:{label}"] = ps
        if residual:
            ps = _plane_stats(residual, code)
            if ps is not None:
                landing[f"{code}:·" if by_plane else code] = ps
    return Stats(n_total=n, fate_counts=fate_counts,
                 transmitted_fate=transmitted_fate, transmission=transmission,
                 per_mz=per_mz, end_spread=spread(ok_rows), r_end_mean=r_mu,
                 r_end_sigma=r_sd, n_col_mean=nc_mu,
                 impact_spread=spread(impact_rows), n_transmitted=len(ok),
                 landing=landing)
