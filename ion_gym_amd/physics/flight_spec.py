"""The deck fields the flight path reads, as plain dataclasses.

A subset of ion_gym's SimSpec (ion_gym/io/sim_spec.py: RFGroupSpec
:591-692, GeometrySpec.field_method/channel_dtype :1339-1340, SourceSpec
:1484, CollisionSpec :1767, IntegrationSpec :2041, BoundsSpec :2230-2262)
with the same names, defaults and units: enough to describe a flight of an
already-built field model.  Two derived values follow the reference's
rules: an unset step (``dt_ns = None``) is the fastest active drive
period / 250, 1 ns without drives (sim_spec.py:99, 2622-2631), and the
field storage precision follows ``geometry.channel_dtype``.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from .ccs import LEGACY_FIXED_SIGMA_M2, SIGMA_MODEL_DEFAULT

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


# the stock step per drive period of a derived dt (sim_spec.py:99)
RF_RECOMMENDED_STEPS_PER_PERIOD = 250
RF_WAVEFORMS = ("sin", "square", "table")


@dataclass
class RFGroupSpec:
    """A named drive group (sim_spec.py:591-692): every member electrode
    carries V(t) = amplitude_v * w(t) + offset_v on top of its dc, with w
    sin(2 pi f t + phase), the duty square (sign(sin) at duty 0.5) or a
    table of dimensionless multipliers (hold or linear, clamped at its
    ends), on the lab clock.  pe_mode concerns only the reference's
    potential-surface display and is carried for decks that state it."""
    name: str
    frequency_hz: float = 5e5
    amplitude_v: float = 50.0
    phase_deg: float = 0.0
    duty: float = 0.5
    offset_v: float = 0.0
    waveform: str = "sin"
    table_t_us: list = field(default_factory=list)
    table_v: list = field(default_factory=list)
    interp: str = "hold"
    pe_mode: Optional[str] = None

    def validate(self):
        """Raise ValueError for a group the flight cannot mean (the
        reference's checks: duty in (0, 1), a known waveform, a table with
        matching non-empty strictly increasing breakpoints and a known
        interpolation)."""
        if not 0.0 < float(self.duty) < 1.0:
            raise ValueError(f"RF group {self.name!r}: duty {self.duty} "
                             f"outside (0, 1)")
        if self.waveform not in RF_WAVEFORMS:
            raise ValueError(f"RF group {self.name!r}: unknown waveform "
                             f"{self.waveform!r}")
        if self.waveform != "table":
            return
        t, v = list(self.table_t_us), list(self.table_v)
        if not t or len(t) != len(v):
            raise ValueError(f"RF group {self.name!r}: table needs matching, "
                             f"non-empty table_t_us/table_v")
        if any(b <= a for a, b in zip(t, t[1:])):
            raise ValueError(f"RF group {self.name!r}: table_t_us must be "
                             f"strictly increasing")
        if self.interp not in ("hold", "linear"):
            raise ValueError(f"RF group {self.name!r}: interp must be "
                             f"'hold' or 'linear'")

    def resolved_pe_mode(self):
        if self.pe_mode in ("pseudo", "instant"):
            return self.pe_mode
        return "pseudo" if self.waveform == "sin" else "instant"


@dataclass
class ElectrodeSpec:
    """A named conductor (sim_spec.py:1000-1036): its metal (``shapes``, a
    list of JSON shape dicts, or an ``stl`` file) and its voltages -- a
    ``dc`` plus membership in any number of drive groups (``rf_groups``)
    and at most one DC ladder (``dc_group`` with its rung ``dc_index`` or
    tap ``dc_weight``).  ``basis`` is the index of its solved basis
    (default: its 1-based position)."""
    name: str
    shapes: List[dict] = field(default_factory=list)
    stl: Optional[str] = None
    is_grid: bool = False
    dc: float = 0.0
    rf_groups: List[str] = field(default_factory=list)
    dc_group: Optional[str] = None
    dc_index: Optional[int] = None
    dc_weight: Optional[float] = None
    color: List[int] = field(default_factory=lambda: [200, 60, 60])
    basis: Optional[int] = None

    def group_names(self):
        """Drive-group memberships, first occurrence order, no repeats."""
        return list(dict.fromkeys(g for g in self.rf_groups if g))


DC_INTERPS = ("linear", "weights")


@dataclass
class DCGroupSpec:
    """A DC ladder fed from two ends (sim_spec.py:901-958): v_in on the
    member with the lowest dc_index, v_out on the highest, the rest
    equally spaced by index ('linear') or placed by their dc_weight
    ('weights'); ``uniform`` puts every member at v_in.  mode 'set' writes
    the member dc, 'add' rides on top of the authored dc."""
    name: str = "LADDER"
    v_in: float = 0.0
    v_out: float = 0.0
    interp: str = "linear"
    mode: str = "set"
    uniform: bool = False


def dc_group_member_values(grp, members):
    """{member name: the group's voltage for it} (sim_spec.py:961-996); a
    weighted ladder with an unweighted member, or an indexed ladder with
    an unnumbered one, is refused by name."""
    v_in, v_out = float(grp.v_in), float(grp.v_out)
    if grp.uniform:
        return {e.name: v_in for e in members}
    if grp.interp == "weights":
        pos = {e.name: e.dc_weight for e in members}
        missing = [n for n, w in pos.items() if w is None]
        if missing:
            raise ValueError(f"dc_group {grp.name!r} (interp='weights'): "
                             f"member(s) {missing} have no dc_weight -- a "
                             f"weighted ladder needs every tap placed")
        return {n: v_in + (v_out - v_in) * float(w) for n, w in pos.items()}
    rung = {e.name: e.dc_index for e in members}
    missing = [n for n, k in rung.items() if k is None]
    if missing:
        raise ValueError(f"dc_group {grp.name!r} (interp={grp.interp!r}): "
                         f"member(s) {missing} have no dc_index -- an "
                         f"indexed ladder needs every rung numbered")
    lo = min(rung.values())
    span = (max(rung.values()) - lo) or 1
    # (v_out - v_in) * (k - lo) first, then / span: the rungs of an
    # integer-volt ladder stay exact
    return {n: v_in + (v_out - v_in) * (k - lo) / span
            for n, k in rung.items()}


FIELD_METHODS = ("electrode_aware", "plain_gradient")
CHANNEL_DTYPES = ("float64", "float32")


@dataclass
class GeometrySpec:
    """The geometry fields the flight reads (sim_spec.py:1252-1345):
    field_method (the 3-D field build: one-sided differences at metal
    nodes, or np.gradient everywhere), channel_dtype (storage of the
    channel stacks: float64, or float32 read back as float64) and the
    drive groups; plus, for a deck whose bases are cached, the solve
    domain (width/height/depth mm, mm_per_gu, symmetry), the electrodes
    and the DC ladders."""
    field_method: str = "electrode_aware"
    channel_dtype: str = "float64"
    rf_groups: List[RFGroupSpec] = field(default_factory=list)
    width_mm: Optional[float] = None
    height_mm: Optional[float] = None
    depth_mm: float = 0.0
    mm_per_gu: float = 0.1
    symmetry: object = "planar"
    electrodes: List[ElectrodeSpec] = field(default_factory=list)
    dc_groups: List[DCGroupSpec] = field(default_factory=list)

    def validate(self):
        if self.field_method not in FIELD_METHODS:
            raise ValueError(f"geometry.field_method {self.field_method!r}: "
                             f"choose one of {FIELD_METHODS}")
        if self.channel_dtype not in CHANNEL_DTYPES:
            raise ValueError(f"geometry.channel_dtype "
                             f"{self.channel_dtype!r}: choose one of "
                             f"{CHANNEL_DTYPES}")
        for g in self.rf_groups:
            g.validate()
        for d in self.dc_groups:
            if d.interp not in DC_INTERPS:
                raise ValueError(f"dc_group {d.name!r}: interp must be one "
                                 f"of {DC_INTERPS}")
            if d.mode not in ("set", "add"):
                raise ValueError(f"dc_group {d.name!r}: mode must be 'set' "
                                 f"or 'add'")


@dataclass
class CollisionSpec:
    """Buffer gas (sim_spec.py:1767-1830).  ``enabled`` False is vacuum.
    Pressure is Torr (P_pa derived).  Hard sphere: the per-ion cross
    section comes from ``sigma_model`` (ion_gym_amd.physics.ccs: 'ccs_fit'
    by default, 'fixed' flies sigma_m2).  SDS: ``gas_diam_nm`` and
    ``sds_species`` (declared K0 or diameter per ion mass; other masses
    are estimated from mass, ion_gym_amd.physics.sds)."""
    enabled: bool = True
    model: str = "hs"            # "hs" | "sds"
    gas: str = "N2"
    T_k: float = 273.0
    P_torr: float = 1.0
    sigma_m2: float = LEGACY_FIXED_SIGMA_M2
    gas_diam_nm: float = 0.366
    flow_mm_us: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    sigma_model: str = SIGMA_MODEL_DEFAULT
    ccs_fit_class: str = "all"
    sds_species: List[dict] = field(default_factory=list)

    @property
    def P_pa(self):
        return self.P_torr * 101325.0 / 760.0


@dataclass
class IntegrationSpec:
    """The step and the recording (sim_spec.py:2041-2110).  ``dt_ns =
    None`` asks for the step derived from the drives
    (FlightSpec.derive_dt_ns)."""
    dt_ns: Optional[float] = 1.0
    t_max_us: float = 400.0
    rec_every: int = 80
    max_records: int = 100000    # 3-D record buffer rows per ion
    record_channels: List[str] = field(
        default_factory=lambda: ["speed", "ke_ev", "e_field"])


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


# engine field storage per channel_dtype: the reference's float32 channels
# are stored in FP32 and read back as FP64 (IG_PREC_F32_STORE)
_STORAGE = {"float64": "f64", "float32": "f32_store"}
PRECISIONS = ("f64", "f32_store", "f32")


@dataclass
class FlightSpec:
    source: SourceSpec = field(default_factory=SourceSpec)
    collisions: Optional[CollisionSpec] = None
    integration: IntegrationSpec = field(default_factory=IntegrationSpec)
    bounds: BoundsSpec = field(default_factory=BoundsSpec)
    geometry: GeometrySpec = field(default_factory=GeometrySpec)
    # station planes: (axis 0/1/2, value_mm, sign, (a_lo, a_hi, b_lo, b_hi),
    # kind 5 impact plate / 6 detector)
    planes: List[tuple] = field(default_factory=list)
    # engine precision override: None follows geometry.channel_dtype;
    # "f32" selects the FP32-basis engine (FP32 fields and interpolation,
    # FP64 state)
    precision: Optional[str] = None
    # the 3-D periodic transporter (sim_spec transporter; tracer3d.py:
    # 722-735): dict(axis, accept_mm, emit_mm, direction, max_passes)
    transporter: Optional[dict] = None
    # declared stations (ion_gym_amd.physics.stations.StationSpec): the
    # kernel-reaching ones become planes before ``planes``
    stations: List = field(default_factory=list)
    # the deck's name (provenance of cached bases) and its builder (part of
    # the basis key; "" dispatches on the geometry)
    name: str = "simulation"
    builder: str = ""

    def resolve_dc_groups(self):
        """Write every 'set'-mode ladder member's dc from its group
        (sim_spec.py:2633-2663); a uniform group's v_out follows v_in.
        Returns the spec."""
        g = self.geometry
        for grp in g.dc_groups:
            members = [e for e in g.electrodes if e.dc_group == grp.name]
            if not members or grp.mode == "add":
                continue
            values = dc_group_member_values(grp, members)
            if grp.uniform:
                grp.v_out = grp.v_in
            for e in members:
                e.dc = values[e.name]
        return self

    def dc_effective(self, el):
        """The static volts electrode ``el`` contributes: its dc, plus its
        'add'-mode group's value for it (sim_spec.py:2665-2686)."""
        base = float(el.dc or 0.0)
        grp = next((x for x in self.geometry.dc_groups
                    if el.dc_group is not None and x.name == el.dc_group),
                   None)
        if grp is None or grp.mode != "add":
            return base
        members = [e for e in self.geometry.electrodes
                   if e.dc_group == grp.name]
        return base + dc_group_member_values(grp, members)[el.name]

    def engine_precision(self):
        """The engine's field precision for this deck."""
        if self.precision is not None:
            if self.precision not in PRECISIONS:
                raise ValueError(f"precision {self.precision!r}: choose one "
                                 f"of {PRECISIONS}")
            return self.precision
        self.geometry.validate()
        return _STORAGE[self.geometry.channel_dtype]

    def derive_dt_ns(self):
        """The step a deck without one flies: the fastest drive period
        with a nonzero frequency and amplitude / 250, or 1 ns for a
        drive-less (DC) deck (sim_spec.py:2622-2631)."""
        fs = [float(g.frequency_hz) for g in self.geometry.rf_groups
              if g.frequency_hz and g.amplitude_v]
        if not fs:
            return 1.0
        return 1e9 / (max(fs) * RF_RECOMMENDED_STEPS_PER_PERIOD)

    def dt_ns(self):
        """The step flown, ns: the stated one, or the derived one for
        ``integration.dt_ns = None`` (refused with a table drive, whose
        breakpoints no period describes, sim_spec.py:3140-3146)."""
        dt = self.integration.dt_ns
        if dt is not None:
            dt = float(dt)
            if not (math.isfinite(dt) and dt > 0.0):
                raise ValueError(f"integration.dt_ns = {dt!r}: the step "
                                 f"must be a finite number > 0")
            return dt
        if any(g.waveform == "table" for g in self.geometry.rf_groups):
            raise ValueError(
                "integration: dt_ns is unset with a table-waveform drive; a "
                "table's breakpoints set no period to derive a step from -- "
                "pin dt_ns explicitly")
        return self.derive_dt_ns()
