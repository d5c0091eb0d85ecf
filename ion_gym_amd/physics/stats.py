"""Ensemble statistics of a flight (ion_gym/physics/stats.py:45-447).

``compute_stats(results, transmitted_fate=...)`` reduces an ensemble's
summaries to fate counts and transmission (the success fate is a required
argument; ``auto_transmitted_fate(spec)`` answers it for transmission
devices), per-mass time-of-flight mean, sigma, FWHM and resolving power of
the transmitted ions, end-point spreads, the mean collision count, and
where ions landed: grouped by declared plane when the plane set is given
(``enabled_planes(spec)``, whose normal is then known), else one group per
fate with the axis of least spread as its normal.  Standard deviations are
population ones (ddof = 0), as the reference states.

The summaries are first laid out as columns (one float array per key,
NaN where a summary lacks the key or holds None), so every reduction below
is an array operation over a boolean row mask.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .ion_envelope import FATE_NAME

FWHM_PER_SIGMA = 2.35482004503
_AXES = "xyz"
_KEYS = ("kind", "tof", "mz", "x_end", "y_end", "z_end", "r_end", "n_col",
         "ke_end", "vx_end", "vy_end", "vz_end")


def auto_transmitted_fate(spec):
    """The fate that counts as transmitted for a transmission device: 3
    (a bounding plane) when the deck enables any bound, else 1 (left the
    domain) (stats.py:69-89)."""
    b = getattr(spec, "bounds", None)
    if b is not None and any(getattr(b, f"{a}_{side}_on", False)
                             for a in _AXES for side in ("min", "max")):
        return 3
    return 1


def mz_of_results(spec, n):
    """Per-ion mass of n results: the source's masses in contiguous blocks
    of n_ions (stats.py:92-108)."""
    masses = list(spec.source.mz_list or [])
    if not masses:
        return [None] * n
    block = max(1, int(getattr(spec.source, "n_ions", 1)))
    return [masses[(i // block) % len(masses)] for i in range(n)]


def enabled_planes(spec):
    """(label, axis, coord_mm) of each enabled bounding plane
    (stats.py:186-201)."""
    b = getattr(spec, "bounds", None) if spec is not None else None
    if b is None:
        return []
    planes = []
    for a in _AXES:
        for side in ("min", "max"):
            if getattr(b, f"{a}_{side}_on", False):
                c = float(getattr(b, f"{a}_{side}"))
                planes.append((f"{a}={c:+g}", a, c))
    return planes


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


class _Columns:
    """The summaries as float columns (NaN for a missing or None value)
    plus the raw fate of each row (fates are compared as given)."""

    def __init__(self, summaries, mz):
        self.n = len(summaries)
        self.kind = [s.get("kind") for s in summaries]
        self.col = {}
        for key in _KEYS[1:]:
            self.col[key] = np.array(
                [np.nan if s.get(key) is None else float(s.get(key))
                 for s in summaries], np.float64)
        # the per-mass grouping key (the caller's masses; None keeps
        # unknown masses apart from numbers)
        self.mz_key = (list(mz) if mz is not None else [None] * self.n)

    def rows(self, mask):
        return np.flatnonzero(mask)

    def fate_is(self, code):
        return np.array([k == code for k in self.kind], bool)


def _moments(v):
    """(mean, population sigma, count) of the finite entries, or
    (None, None, 0) when there are none."""
    v = v[np.isfinite(v)]
    if v.size == 0:
        return None, None, 0
    return float(v.mean()), float(v.std()) if v.size > 1 else 0.0, int(v.size)


def _tof_width(tof, masses):
    """Arrival-time block of a landing group (stats.py:204-223): blank
    with a stated reason for mixed masses or no recorded tof, and no width
    below two ions (or with none)."""
    t = tof[np.isfinite(tof)]
    if t.size == 0:
        return dict(tof_n=0, tof_note="route records no tof")
    if len({m for m in masses if m is not None}) > 1:
        return dict(tof_n=0, tof_note="mixed masses — see the per-mass tabs")
    mu, sd, cnt = _moments(t)
    if cnt < 2 or not sd:
        return dict(tof_n=cnt, tof_mean=mu, tof_note="n < 2 — no width")
    w = FWHM_PER_SIGMA * sd
    return dict(tof_n=cnt, tof_mean=mu, tof_fwhm=w, tof_R=mu / (2 * w))


def _landing(cols, idx, fate, normal=None, label=""):
    """Where the rows idx of one fate landed (stats.py:226-297): spread
    about the centroid in the plane across ``normal`` (the axis of least
    spread when not given), energy, velocity split and divergence, and the
    arrival-time block.  None for an empty group or one missing an end
    point."""
    if idx.size == 0:
        return None
    end = np.stack([cols.col[f"{a}_end"][idx] for a in _AXES], axis=1)
    if not np.all(np.isfinite(end[:, 0])):
        return None
    mean = end.mean(axis=0)
    sig = end.std(axis=0) if idx.size > 1 else np.zeros(3)
    if normal is None or normal not in _AXES:
        normal = _AXES[int(np.argmin(sig))]
    k = _AXES.index(normal)
    across = [q for q in range(3) if q != k]
    c0, c1 = float(mean[across[0]]), float(mean[across[1]])
    r = np.sort(np.hypot(end[:, across[0]] - c0, end[:, across[1]] - c1))
    n = r.size
    ke = cols.col["ke_end"][idx]
    ke = ke[np.isfinite(ke)]
    ke_mu, ke_sd, _ = _moments(ke)
    v = np.stack([cols.col[f"v{a}_end"][idx] for a in _AXES], axis=1)
    v = v[np.all(np.isfinite(v), axis=1)]
    vel = dict(v_norm_mean=0.0, v_norm_sigma=0.0, v_perp_mean=0.0,
               v_perp_sigma=0.0, v_perp_rms=0.0, div_mrad_rms=0.0,
               div_mrad_p95=0.0)
    if v.shape[0]:
        vn = v[:, k]
        vp = np.hypot(v[:, across[0]], v[:, across[1]])
        several = v.shape[0] > 1
        vel.update(v_norm_mean=float(vn.mean()),
                   v_norm_sigma=float(vn.std()) if several else 0.0,
                   v_perp_mean=float(vp.mean()),
                   v_perp_sigma=float(vp.std()) if several else 0.0,
                   v_perp_rms=float(np.sqrt(np.mean(vp * vp))))
        with np.errstate(invalid="ignore", divide="ignore"):
            mrad = 1e3 * np.abs(np.arctan2(vp, np.abs(vn)))
        mrad = mrad[np.isfinite(mrad)]
        if mrad.size:
            vel.update(div_mrad_rms=float(np.sqrt(np.mean(mrad * mrad))),
                       div_mrad_p95=float(np.percentile(mrad, 95.0)))
    # mixed masses are told by the summaries' own mz
    own = cols.col["mz"][idx]
    masses = list(own[np.isfinite(own)])
    return PlaneStats(
        fate=fate, n=n, normal=normal, normal_mean=float(mean[k]),
        normal_sigma=float(sig[k]), in_plane=tuple(_AXES[q] for q in across),
        centroid=(c0, c1), sigma=(float(sig[across[0]]),
                                  float(sig[across[1]])),
        r_mean=float(r.mean()), r_rms=float(np.sqrt(np.mean(r * r))),
        r_p95=float(r[min(n - 1, int(0.95 * n))]), plane_label=label,
        ke_n=int(ke.size), ke_mean=ke_mu or 0.0, ke_sigma=ke_sd or 0.0,
        ke_min=float(ke.min()) if ke.size else 0.0,
        ke_max=float(ke.max()) if ke.size else 0.0, **vel,
        **_tof_width(cols.col["tof"][idx], masses))


def _spread(cols, idx):
    out = {}
    for a in _AXES:
        mu, sd, cnt = _moments(cols.col[f"{a}_end"][idx])
        if cnt:
            out[a] = (mu, sd)
    return out


def _plane_tolerance(planes, tol):
    """The match tolerance: at most half the smallest gap between two
    declared planes on one axis."""
    for a in _AXES:
        cs = np.unique([c for _, ax, c in planes if ax == a])
        if cs.size > 1:
            tol = min(tol, 0.5 * float(np.diff(cs).min()))
    return tol


def _assign_planes(cols, idx, planes, tol):
    """For rows idx, the index of the nearest declared plane within tol
    (distance along the plane's own axis), or -1."""
    dist = np.full((idx.size, len(planes)), np.inf)
    for p, (_, a, c) in enumerate(planes):
        d = np.abs(cols.col[f"{a}_end"][idx] - c)
        dist[:, p] = np.where(np.isfinite(d), d, np.inf)
    best = np.argmin(dist, axis=1) if planes else np.zeros(idx.size, int)
    ok = dist[np.arange(idx.size), best] <= tol if planes else \
        np.zeros(idx.size, bool)
    return np.where(ok, best, -1)


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
    summaries = [r.summary if hasattr(r, "summary") else r for r in results]
    cols = _Columns(summaries, mz)
    n = cols.n
    fate_counts = {}
    for k in cols.kind:
        fate_counts[k] = fate_counts.get(k, 0) + 1
    ok = cols.rows(cols.fate_is(transmitted_fate))

    # per mass, over the transmitted ions (masses sorted, None last)
    per_mz = []
    ok_masses = [cols.mz_key[i] for i in ok]
    for m in sorted(set(ok_masses), key=lambda v: (v is None, v)):
        sel = ok[[q for q, mm in enumerate(ok_masses) if mm == m]]
        mu, sd, cnt = _moments(cols.col["tof"][sel])
        fwhm = rp = None
        if mu is not None:
            fwhm = FWHM_PER_SIGMA * sd
            rp = mu / (2 * fwhm) if fwhm > 0 else float("inf")
        per_mz.append(MzStats(m, cnt, mu, sd, fwhm, rp))

    r_mu, r_sd, _ = _moments(cols.col["r_end"][ok])
    nc_mu, _, _ = _moments(cols.col["n_col"])

    landing = {}
    plist = list(planes) if planes else []
    tol = _plane_tolerance(plist, float(plane_tol_mm))
    for code in sorted(fate_counts, key=lambda c: (c is None, c)):
        idx = cols.rows(cols.fate_is(code))
        if not plist:
            ps = _landing(cols, idx, code)
            if ps is not None:
                landing[code] = ps
            continue
        which = _assign_planes(cols, idx, plist, tol)
        hit = sorted({int(p) for p in which if p >= 0},
                     key=lambda p: (plist[p][0], plist[p][1]))
        for p in hit:
            label, axis, _ = plist[p]
            ps = _landing(cols, idx[which == p], code, normal=axis,
                          label=label)
            if ps is not None:
                landing[f"{code}:{label}"] = ps
        rest = idx[which < 0]
        if rest.size:
            ps = _landing(cols, rest, code)
            if ps is not None:
                landing[f"{code}:·" if hit else code] = ps

    return Stats(n_total=n, fate_counts=fate_counts,
                 transmitted_fate=transmitted_fate,
                 transmission=(ok.size / n) if n else None, per_mz=per_mz,
                 end_spread=_spread(cols, ok), r_end_mean=r_mu,
                 r_end_sigma=r_sd, n_col_mean=nc_mu,
                 impact_spread=_spread(cols, cols.rows(cols.fate_is(0))),
                 n_transmitted=int(ok.size), landing=landing)
