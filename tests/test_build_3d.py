"""The 3-D fly function's host parts without a device: the kernel plane
list (stations, then bounds as fate-3 whole-plane kills), the 3-D birth
refusal against the oracle's restatement, the escape face and the edge
field guard, and the refusals made before any flight."""
import math

import numpy as np
import pytest

from oracle import births as OB
from ion_gym_amd.physics import build_3d as B3
from ion_gym_amd.physics.drift_extension import (xyz_edge_fields,
                                                 xyz_escape_face)
from ion_gym_amd.physics.flight_spec import (BoundsSpec, CollisionSpec,
                                             FlightSpec, IntegrationSpec)
from ion_gym_amd.physics.ion_envelope import (BirthInMetal,
                                              BirthOutsideDomain,
                                              xyz_birth_refusal)


def test_kernel_planes_stations_then_bounds():
    st = (2, 25.0, 0.0, (0.5, 3.5, 0.5, 3.5), 6)
    spec = FlightSpec(planes=[st], bounds=BoundsSpec(x_min_on=True,
                                                     x_min=-1.0,
                                                     z_max_on=True,
                                                     z_max=30.0))
    pl = B3.kernel_planes(spec)
    assert pl[0] == st
    assert pl[1][:3] == (0, -1.0, -1.0) and pl[1][4] == 3
    assert pl[2][:3] == (2, 30.0, 1.0) and pl[2][4] == 3
    w = pl[1][3]
    assert w[0] > w[1] and w[2] > w[3]     # no point is inside


def test_xyz_birth_refusal_matches_oracle():
    rng = np.random.default_rng(3)
    m = np.zeros((12, 10, 14), bool)
    m[3:7, 2:6, 5:11] = True
    m[9, 8, 2] = True
    h = 0.25
    n = 3000
    ext = [(s - 1) * h for s in m.shape]
    x, y, z = (rng.uniform(-0.3, e + 0.3, n) for e in ext)
    want = OB.refuse_3d(m, h, x, y, z)
    got = []
    for i in range(n):
        e = xyz_birth_refusal(m, h, x[i], y[i], z[i], i)
        got.append(-1 if e is None else
                   (1 if isinstance(e, BirthOutsideDomain) else 0))
        if e is not None:
            assert isinstance(e, BirthInMetal)
    assert np.array_equal(np.array(got), want)
    assert (want == 0).sum() > 30 and (want == 1).sum() > 30


def test_escape_face_and_edge_guard():
    shape, h = (11, 11, 41), 0.5
    assert xyz_escape_face(shape, h, (2.0, 3.0, 20.3))[:3] == ("z_hi", 2,
                                                               "hi")
    assert xyz_escape_face(shape, h, (-0.1, 3.0, 5.0))[:3] == ("x_lo", 0,
                                                               "lo")
    with pytest.raises(ValueError):
        xyz_escape_face(shape, h, (1.0, 1.0, 1.0))
    z = np.zeros(shape)
    pack = dict(EAx=z, EAy=z, EAz=z.copy(), ExK=np.zeros((1,) + shape),
                EyK=np.zeros((1,) + shape), EzK=np.zeros((1,) + shape),
                ch_kind=np.array([0]), ch_amp=np.array([3.0]),
                ch_off=np.array([0.5]))
    # fields on the interiors of the z faces only (the x and y faces
    # share their edge nodes)
    pack["EzK"][0, 1:-1, 1:-1, -1] = 2.0
    pack["EAz"][1:-1, 1:-1, 0] = 1.0
    e = xyz_edge_fields(pack)
    # the drive channel's worst weight is |amp| + |off|
    assert math.isclose(e["z_hi"], 2.0 * 3.5)
    assert math.isclose(e["z_lo"], 1.0) and e["x_lo"] == 0.0


def test_refusals_before_flight():
    pack = dict(ele=np.zeros((3, 3, 3), bool), h_mm=1.0)
    births = np.zeros((1, 7))
    with pytest.raises(ValueError, match="cannot record"):
        B3.make_3d_fly_fn(pack, births, FlightSpec(
            integration=IntegrationSpec(record_channels=["path_mm"])))
    spec = FlightSpec()
    spec.source.charge = 0
    with pytest.raises(ValueError, match="charge = 0"):
        B3.make_3d_fly_fn(pack, births, spec)
    spec = FlightSpec(collisions=CollisionSpec(enabled=True, model="sds"),
                      integration=IntegrationSpec(
                          record_channels=["wrap_passes"]))
    with pytest.raises(ValueError, match="wrap_passes"):
        B3.make_3d_fly_fn(pack, births, spec)
    spec = FlightSpec(collisions=CollisionSpec(enabled=True, model="ems"))
    with pytest.raises(ValueError, match="is not a known model"):
        B3.make_3d_fly_fn(pack, births, spec)
