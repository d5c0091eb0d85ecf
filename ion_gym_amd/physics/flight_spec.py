"""The deck fields the flight path reads, as plain dataclasses.

A subset of ion_gym's SimSpec (ion_gym/io/sim_spec.py: SourceSpec :1484,
CollisionSpec :1767, IntegrationSpec :2041, BoundsSpec :2230-2262) with the
same names, defaults where the reference has one, and units: enough to
describe a flight of an already-built field model.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple


@dataclass
class SourceSpec:
    n_ions: int = 100
    mz_list: List[float] = field(default_factory=lambda: [100.0])
    charge: int = 1
    seed: Optional[int] = 0


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
