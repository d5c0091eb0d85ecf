"""The deck fields the flight path reads, as plain dataclasses.

A subset of ion_gym's SimSpec (ion_gym/io/sim_spec.py: SourceSpec :1484,
CollisionSpec :1767, IntegrationSpec :2041, BoundsSpec :2230-2262) with the
same names, defaults where the reference has one, and units: enough to
describe a flight of an already-built field model.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

FWHM_PER_SIGMA = 2.0 * math.sqrt(2.0 * math.log(2.0))
SOURCE_DISTRIBUTIONS = ("point", "disc", "line", "grid", "box", "gaussian",
                        "file")


@dataclass
class SourceSpec:
    """Ion births (ion_gym/io/sim_spec.py:1484-1720): where
    (distribution about (x0, y0, z0): point, disc of radius r_mm normal
    to ``axis``, line of len_mm along ``axis``, grid lattice on the disc,
    box of full extents box_mm, truncated gaussian with per-axis fwhm_mm
    and trunc_mm half-widths, or a births_file CSV), how fast (ke_lo..ke_hi
    eV along ``direction``, plus a temperature_k thermal spread, plus
    per-axis dv_fwhm_ms spreads in m/s, plus a v_drift_mm_us drift) and
    when (tob uniform over tob_span_us).  n_ions is per mass; masses
    (Da) are flown in contiguous blocks."""
    n_ions: int = 100
    mz_list: List[float] = field(default_factory=lambda: [100.0])
    charge: int = 1
    seed: Optional[int] = 0
    distribution: str = "disc"
    x0_mm: float = 1.0
    y0_mm: float = 0.0
    z0_mm: float = 0.0
    r_mm: float = 4.0
    len_mm: float = 0.0
    box_mm: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])
    axis: str = "x"
    direction: List[float] = field(default_factory=lambda: [1.0, 0.0, 0.0])
    ke_lo: float = 0.1
    ke_hi: float = 1.9
    temperature_k: float = 0.0
    dv_fwhm_ms: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    fwhm_mm: Optional[List[float]] = None
    trunc_mm: Optional[List[float]] = None
    v_drift_mm_us: List[float] = field(
        default_factory=lambda: [0.0, 0.0, 0.0])
    tob_span_us: float = 2.0
    births_file: Optional[str] = None

    def gaussian_layout(self):
        """(sigmas, truncations) in mm per lab axis for the gaussian
        distribution (sim_spec.py:1583-1628).  Refused with ValueError:
        missing or non-3-vector fwhm_mm/trunc_mm, negative entries, all
        widths zero, a truncation below one sigma, a truncation on an axis
        without spread."""
        f, t = self.fwhm_mm, self.trunc_mm
        for name, v in (("fwhm_mm", f), ("trunc_mm", t)):
            if v is None:
                raise ValueError(
                    f"source: distribution='gaussian' requires {name} (a "
                    f"3-vector, mm); no beam width is assumed")
            if not hasattr(v, "__len__") or len(v) != 3:
                raise ValueError(f"source: {name} must be a 3-vector "
                                 f"(x, y, z) in mm, got {v!r}")
        f = [float(x) for x in f]
        t = [float(x) for x in t]
        if any(x < 0 for x in f + t):
            raise ValueError(f"source: fwhm_mm/trunc_mm entries must be >= "
                             f"0, got fwhm_mm={f}, trunc_mm={t}")
        if not any(x > 0 for x in f):
            raise ValueError("source: all three fwhm_mm entries are 0: a "
                             "point source; declare distribution='point'")
        sig = [x / FWHM_PER_SIGMA for x in f]
        for k, ax in enumerate("xyz"):
            if f[k] > 0 and t[k] < sig[k]:
                raise ValueError(
                    f"source: trunc_mm[{ax}] = {t[k]:g} mm is below 1 sigma "
                    f"({sig[k]:.4g} mm) of fwhm_mm[{ax}] = {f[k]:g} mm; "
                    f"declare such a beam as a box")
            if f[k] == 0 and t[k] != 0:
                raise ValueError(
                    f"source: trunc_mm[{ax}] = {t[k]:g} mm on an axis with "
                    f"fwhm_mm[{ax}] = 0; set it to 0")
        return sig, t

    def direction_error(self):
        """None, or why ``direction`` is not a 3-vector of finite floats
        (sim_spec.py:1690-1703)."""
        d = self.direction
        try:
            ok = len(d) == 3 and all(math.isfinite(float(v)) for v in d)
        except (TypeError, ValueError):
            ok = False
        if ok:
            return None
        return (f"source.direction must be a 3-vector of finite floats, "
                f"got {d!r}")

    def zero_direction_note(self):
        """The warning for a beam energy declared with a zero direction
        (sim_spec.py:1705-1726): the energy is still drawn but ignored, so
        the ions are born at rest plus any thermal or per-axis spread."""
        err = self.direction_error()
        if err:
            raise ValueError(err)
        zero = not any(float(v) != 0.0 for v in self.direction)
        if not zero or not (self.ke_lo > 0 or self.ke_hi > 0):
            return None
        return (f"WARNING: source.direction is the zero vector, so "
                f"ke_lo..ke_hi = {self.ke_lo:g}..{self.ke_hi:g} eV has "
                f"nowhere to point and is IGNORED: ions are born at rest")


@dataclass
class CollisionSpec:
    enabled: bool = False
    model: str = "hs"            # "hs" | "sds"
    gas: str = "N2"
    T_k: float = 273.0
    P_torr: float = 1.0
    sigma_m2: float = 1.0e-18
    flow_mm_us: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    K0_cm2_per_Vs: Optional[float] = None   # SDS: declared mobility

    @property
    def P_pa(self):
        return self.P_torr * 101325.0 / 760.0


@dataclass
class IntegrationSpec:
    dt_ns: float = 1.0
    t_max_us: float = 50.0
    rec_every: int = 10
    record_channels: List[str] = field(default_factory=list)
    max_records: int = 200000    # 3-D record buffer rows per ion


@dataclass
class BoundsSpec:
    x_min_on: bool = False
    x_max_on: bool = False
    y_min_on: bool = False
    y_max_on: bool = False
    z_min_on: bool = False
    z_max_on: bool = False
    x_min: float = 0.0
    x_max: float = 0.0
    y_min: float = 0.0
    y_max: float = 0.0
    z_min: float = 0.0
    z_max: float = 0.0

    def as_tuple(self):
        return ((self.x_min_on, self.x_max_on, self.y_min_on,
                 self.y_max_on, self.z_min_on, self.z_max_on),
                (self.x_min, self.x_max, self.y_min, self.y_max,
                 self.z_min, self.z_max))


@dataclass
class FlightSpec:
    source: SourceSpec = field(default_factory=SourceSpec)
    collisions: Optional[CollisionSpec] = None
    integration: IntegrationSpec = field(default_factory=IntegrationSpec)
    bounds: BoundsSpec = field(default_factory=BoundsSpec)
    # station planes: (axis 0/1/2, value_mm, sign, (a_lo, a_hi, b_lo, b_hi),
    # kind 5 impact plate / 6 detector)
    planes: List[tuple] = field(default_factory=list)
    precision: str = "f64"       # engine field storage: f64 | f32_store | f32
    # the 3-D periodic transporter (sim_spec transporter; tracer3d.py:
    # 722-735): dict(axis, accept_mm, emit_mm, direction, max_passes)
    transporter: Optional[dict] = None
    # declared stations (ion_gym_amd.physics.stations.StationSpec): the
    # kernel-reaching ones become planes before ``planes``
    stations: List = field(default_factory=list)
