"""The ballistic drift beyond the solved domain (drift_extension.py of the
reference, restated): exact crossing records for a bound and for station
planes, the engagement rule, the refusals and the channel rules."""
import math

import numpy as np
import pytest

from ion_gym_amd.physics import drift_extension as D
from ion_gym_amd.physics.flight_spec import (BoundsSpec, CollisionSpec,
                                             FlightSpec, IntegrationSpec)

COLS = ["t", "x", "y", "z", "vx", "vy", "vz", "speed", "ke_ev", "e_axial",
        "radius", "path_mm", "ke_tint", "n_col"]
# the route supplies the radius rule (build_rz.py:766-768)
RADIUS = {"radius": lambda p, row, idx, dt: math.hypot(row[idx["y"]],
                                                       row[idx["z"]])}


def _traj(x=40.0, y=1.0, z=0.5, v=(2.0, 0.1, -0.05), t=10.0):
    sp = math.sqrt(sum(c * c for c in v))
    row0 = [0.0, 0.0, y, z, *v, sp, 100.0, 5.0, math.hypot(y, z), 0.0, 0.0,
            3]
    row1 = [t, x, y, z, *v, sp, 100.0, 7.0, math.hypot(y, z), 20.0, 900.0,
            3]
    return np.array([row0, row1], float)


def _spec(**kw):
    s = FlightSpec(integration=IntegrationSpec(t_max_us=100.0))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_bound_beyond_the_exit_face_gets_an_exact_record():
    tr = _traj()
    spec = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0))
    out, kind, info = D.extend_ballistic(tr, COLS, 1, spec,
                                         edge_e_vpermm=0.0, face_axis=0,
                                         face_side="hi", derived=RADIUS)
    assert kind == 3 and out.shape == (3, len(COLS))
    dt = (60.0 - 40.0) / 2.0
    last = out[-1]
    assert last[0] == 10.0 + dt and last[1] == 60.0
    assert last[2] == 1.0 + 0.1 * dt and last[3] == 0.5 - 0.05 * dt
    assert last[COLS.index("e_axial")] == 0.0             # declared zero
    assert last[COLS.index("speed")] == tr[-1, COLS.index("speed")]
    sp = tr[-1, COLS.index("speed")]
    assert last[COLS.index("path_mm")] == 20.0 + sp * dt
    assert last[COLS.index("ke_tint")] == 900.0 + 100.0 * dt
    assert info["bnd_face"] == 1 and np.array_equal(tr, _traj())


def test_radius_comes_from_the_route_rule():
    tr = _traj()
    spec = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0))
    out, _, _ = D.extend_ballistic(
        tr, COLS, 1, spec, edge_e_vpermm=0.0, face_axis=0, face_side="hi",
        derived={"radius": lambda p, row, idx, dt:
                 math.hypot(row[idx["y"]], row[idx["z"]])})
    assert out[-1, COLS.index("radius")] == math.hypot(out[-1, 2],
                                                       out[-1, 3])


def test_terminating_planes_win_over_bounds_and_passing_planes_extend():
    tr = _traj()
    det = (0, 50.0, 0.0, (-5.0, 5.0, -5.0, 5.0), 6)       # detector patch
    plate = (0, 45.0, 0.0, (-5.0, 5.0, -5.0, 5.0), 5)     # aperture plate
    spec = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0),
                 planes=[plate, det])
    out, kind, info = D.extend_ballistic(tr, COLS, 1, spec,
                                         edge_e_vpermm=0.0, face_axis=0,
                                         face_side="hi", derived=RADIUS)
    # the plate's aperture passes the ion; the detector patch stops it
    assert kind == 6 and out[-1, 1] == 50.0 and info["bnd_face"] is None
    # only a passing plane ahead: the record sits on it, the fate stays 1
    spec2 = _spec(planes=[plate])
    out2, kind2, _ = D.extend_ballistic(tr, COLS, 1, spec2,
                                        edge_e_vpermm=0.0, face_axis=0,
                                        face_side="hi", derived=RADIUS)
    assert kind2 == 1 and out2[-1, 1] == 45.0


def test_nothing_declared_beyond_the_escape_face_is_left_alone():
    tr = _traj()
    spec = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0))
    # a radial escape with only an axial target: no extension
    out, kind, info = D.extend_ballistic(tr, COLS, 1, spec,
                                         edge_e_vpermm=5.0, face_axis=1,
                                         face_side="hi")
    assert info is None and kind == 1 and out is tr
    # out of time before the bound
    spec2 = _spec(bounds=BoundsSpec(x_max_on=True, x_max=1e6))
    out, kind, info = D.extend_ballistic(tr, COLS, 1, spec2,
                                         edge_e_vpermm=0.0, face_axis=0,
                                         face_side="hi")
    assert info is None
    # other fates are never touched
    assert D.extend_ballistic(tr, COLS, 0, spec, edge_e_vpermm=0.0,
                              face_axis=0, face_side="hi")[2] is None


def test_refusals():
    tr = _traj()
    spec = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0))
    with pytest.raises(ValueError, match="residual field"):
        D.extend_ballistic(tr, COLS, 1, spec, edge_e_vpermm=0.01,
                           face_axis=0, face_side="hi")
    gas = _spec(bounds=BoundsSpec(x_max_on=True, x_max=60.0),
                collisions=CollisionSpec(enabled=True))
    with pytest.raises(ValueError, match="collision-free"):
        D.extend_ballistic(tr, COLS, 1, gas, edge_e_vpermm=0.0,
                           face_axis=0, face_side="hi")
    # radius has no universal rule: the route must supply it
    with pytest.raises(ValueError, match="no field-free extension rule"):
        D.extend_ballistic(tr, COLS, 1, spec, edge_e_vpermm=0.0,
                           face_axis=0, face_side="hi")


def test_escape_faces_and_edge_fields():
    assert D.rz_escape_face(-0.1, 0.1, 401)[:3] == ("x_lo", 0, "lo")
    assert D.rz_escape_face(40.05, 0.1, 401)[:3] == ("x_hi", 0, "hi")
    assert D.rz_escape_face(20.0, 0.1, 401)[:3] == ("r_hi", 1, "hi")

    class M:
        EzA = np.zeros((5, 9))
        EuA = np.zeros((5, 9))
        EzK = np.zeros((1, 5, 9))
        EuK = np.zeros((1, 5, 9))
        EzG = None
        EuG = None
    M.EzA[-1, :] = 3000.0          # V/m on the x_hi face
    M.EuK[0, 1:-1, -1] = 4000.0    # on the radial edge, off the corners
    e = D.rz_edge_fields(M)
    assert e["x_lo"] == 0.0 and e["x_hi"] == pytest.approx(3.0)
    assert e["r_hi"] == pytest.approx(4.0)
