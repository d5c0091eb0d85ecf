"""ion_gym's round-bore einzel gate (tests/validate_einzel_rz.py:1-28,
156-276) on the engine, from the gate's own deck
(tests/fixtures/einzel_rz_gate_fixture.json, kept as data under
tests/golden/einzel_rz/): three coaxial tubes (bore 6 mm, wall 1 mm,
9.5 mm long, 1 mm gaps) on a 52 x 8 mm r-z domain at 0.1 mm, outer tubes
grounded, the centre one at the lens voltage; 12 ions of m/z 100 at
50 eV from a 1.2 mm disc on x = 1.5 mm, vacuum, dt 0.5 ns, a bound at
x = 51 mm.

The deck path is the product's: the tubes painted on the node grid
(configs.rz_deck_labels, the reference's closed-edge rect rule,
raster2d.py:113-137), one basis per electrode at 1e4 V solved on the
device (ig_solve2d, mirror ghosts on the open edges as build_rz.py:375-
378 asks), stored in the basis cache, composed at each lens voltage
(fields_rz.compose_spec_rz), births drawn on the device (ig_births) and
flown by the engine-backed fly function.  The reference's gates:

  RZ-1  on-axis saddle: -100 V on the lens gives -100 < V(axis, lens
        centre) < -60 and |V(axis, z = 5 mm)| < 5 (validate_einzel_rz.py:
        158-170);
  RZ-2  the mean beam radius at z = 36 mm over lens voltages 0, -80,
        -200, -400 V falls below 0.4 of the field-free radius, then rises
        again (over-focus) (:172-186);
  RZ-4  a new lens voltage re-weights the cached bases: no solve, under
        0.5 s (:211-218);
  RZ-5  three coaxial tubes, each longer than its wall, the axis shielded
        below -60 V at -100 V (:220-253);
  RZ-6  at -120 V every ion stops at the x = 51 mm bound (within 3 mm of
        the 52 mm domain end) and the beam is focused below 0.3 of its
        birth radius (:255-272).

RZ-3 compares against the planar slot einzel, a deck of cutout-ellipse
plates that this r-z path does not paint; it is not run here.  On top of
the gates, each flight is checked ion by ion against the oracle flying
the same composed model and the same births: the FP64 bar (1e-9 relative
per record column, identical fates), and for the FP32-basis engine the
same gates.
"""
import json
import math
import os
import time

import numpy as np
import pytest

from oracle import births as OB
from oracle import oracle as O
from ion_gym_amd.engine import RZ_CHANNELS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "einzel_rz",
                       "einzel_rz_gate_fixture.json")
REL = 1e-9


def _deck(lens_v, precision=None):
    """The fixture as a FlightSpec: its r-z rects (x axial, y radial) as
    (z0, z1, r0, r1) rects, its source, integration and bounds."""
    from ion_gym_amd.physics.flight_spec import (
        BoundsSpec, CollisionSpec, ElectrodeSpec, FlightSpec, GeometrySpec,
        IntegrationSpec, SourceSpec)
    with open(FIXTURE, encoding="utf-8") as fh:
        d = json.load(fh)
    g = d["geometry"]
    assert g["symmetry"]["coords"] == "rz"
    els = []
    for e in g["electrodes"]:
        (s,) = e["shapes"]
        assert s["type"] == "rect"
        els.append(ElectrodeSpec(
            name=e["name"], dc=float(e["dc"]),
            shapes=[{"type": "rect", "z0_mm": s["x_mm"],
                     "z1_mm": s["x_mm"] + s["width_mm"],
                     "r0_mm": s["y_mm"], "r1_mm": s["y_mm"] + s["height_mm"]}]))
    geo = GeometrySpec(width_mm=g["width_mm"], height_mm=g["height_mm"],
                       mm_per_gu=g["mm_per_gu"], symmetry="cylindrical",
                       electrodes=els)
    src = d["source"]
    so = SourceSpec(**{k: src[k] for k in (
        "n_ions", "mz_list", "charge", "seed", "distribution", "x0_mm",
        "y0_mm", "z0_mm", "r_mm", "axis", "direction", "ke_lo", "ke_hi",
        "temperature_k", "tob_span_us")})
    it = d["integration"]
    spec = FlightSpec(
        source=so, collisions=CollisionSpec(enabled=False),
        integration=IntegrationSpec(
            dt_ns=it["dt_ns"], t_max_us=it["t_max_us"],
            rec_every=it["rec_every"],
            record_channels=list(it["record_channels"])),
        bounds=BoundsSpec(**d["bounds"]), geometry=geo,
        name="einzel_rz_gate", precision=precision)
    from ion_gym_amd.physics.flight_conditions import set_flight_conditions
    set_flight_conditions(spec, voltages={"lens": float(lens_v)}, show=False)
    return spec


def _edges(spec):
    """Axial edges (mm) of the three tubes, in deck order."""
    out = []
    for e in spec.geometry.electrodes:
        (s,) = e.shapes
        out += [s["z0_mm"], s["z1_mm"]]
    return out


@pytest.fixture(scope="module")
def solved(tmp_path_factory):
    """Labels, device-solved bases (one per tube at 1e4 V) published in a
    basis cache, the device births of the deck's source, the engine."""
    from ion_gym_amd import births as B
    from ion_gym_amd import configs as C
    from ion_gym_amd.io import basis_cache as BC
    from ion_gym_amd.physics import solver2d as S
    from ion_gym_amd.physics.build_rz import shared_engine
    eng = shared_engine(0)
    spec = _deck(-100.0)
    lab = C.rz_deck_labels(spec)
    h = spec.geometry.mm_per_gu
    print(f"einzel gate deck {lab.shape}, device solve of 3 bases",
          flush=True)
    solve = S.laplace_factor(lab > 0, h, symmetry="cylindrical",
                             edge_ghost="mirror", engine=eng)
    keys = sorted(int(v) for v in np.unique(lab) if v)
    bases = {k: solve(np.where(lab == k, 1e4, 0.0)) for k in keys}
    root = tmp_path_factory.mktemp("einzel_bases")
    BC.store(spec, bases, lab, root=root)
    births = B.generate_births(spec.source, eng)
    return dict(spec=spec, lab=lab, bases=bases, root=root, births=births,
                eng=eng)


def _model(solved, lens_v):
    from ion_gym_amd.io import basis_cache as BC
    from ion_gym_amd.physics import fields_rz as FR
    spec = _deck(lens_v)
    got, ele = BC.load(spec, root=solved["root"])
    assert got is not None, "a voltage change must hit the cached bases"
    return FR.compose_spec_rz(spec, got, ele.astype(np.int32))


def _fly(solved, lens_v, precision=None):
    """[(traj, summary)] of the 12 ions at this lens voltage, and the
    composed model."""
    from ion_gym_amd.physics.build_rz import make_rz_fly_fn
    spec = _deck(lens_v, precision)
    model = _model(solved, lens_v)
    births = solved["births"]
    fly_fn, cols = make_rz_fly_fn(model, births, spec)
    print(f"einzel gate flight at {lens_v} V ({precision or 'f64'})",
          flush=True)
    return fly_fn.batch(range(len(births))), cols, model, spec


def _beam_r(results, zq):
    """Mean |r| at the first record with x >= zq (validate_einzel_rz.py:
    144-153), over ions that reach it; NaN when fewer than three do."""
    rs = []
    for tr, _ in results:
        k = np.searchsorted(tr[:, 1], zq)
        if 0 < k < len(tr):
            rs.append(math.hypot(tr[k, 2], tr[k, 3]))
    return float(np.mean(rs)) if len(rs) > 2 else float("nan")


def _oracle(model, spec, births):
    from ion_gym_amd.physics.constants import E_CHG, KG_AMU
    n = len(births)
    m = np.full(n, float(spec.source.mz_list[0]))
    ions = dict(x=births[:, 0], y=births[:, 1], z=births[:, 2],
                vx=births[:, 3], vy=births[:, 4], vz=births[:, 5],
                tob=births[:, 6], m_amu=m,
                acc=spec.source.charge * E_CHG / (m * KG_AMU),
                sigma_m2=np.zeros(n),
                seed=int(spec.source.seed) + np.arange(n, dtype=np.int64))
    dt = spec.integration.dt_ns * 1e-3
    t_max = spec.integration.t_max_us
    rec = spec.integration.rec_every
    on, val = spec.bounds.as_tuple()
    return O.fly_rz(model.pack(),
                    dict(dt=dt, t_max=t_max, rec_every=rec,
                         rec_rows=int(t_max / dt / rec) + 8,
                         flags=[int(c in spec.integration.record_channels)
                                for c in RZ_CHANNELS],
                         bnd_on=[int(b) for b in on], bnd_val=list(val)),
                    ions)


def test_deck_births_and_bases(solved):
    """The device births of the fixture's disc are the oracle's (4e-15 on
    the trigonometric positions, the rest bit for bit), and the device
    bases solve the reference's cylindrical system (the oracle's LU of
    it, oracle/solve2d.py) to 1e-9 of the 1e4 V scale."""
    from oracle import solve2d as OS
    spec, lab, births = solved["spec"], solved["lab"], solved["births"]
    o = OB.generate_births(spec.source, int(spec.source.seed))
    assert births.shape == o.shape == (12, 7)
    assert np.allclose(births, o, rtol=4e-15, atol=4e-15 * 52.0)
    assert np.array_equal(births[:, 3:], o[:, 3:])
    r = np.hypot(births[:, 1], births[:, 2])
    assert np.all(r <= 1.2) and np.all(births[:, 0] == 1.5)
    h = spec.geometry.mm_per_gu
    for k, b in solved["bases"].items():
        ref, _ = OS.solve(lab > 0, np.where(lab == k, 1e4, 0.0), h,
                          "cylindrical", edge_ghost="mirror")
        assert float(np.abs(b - ref).max()) <= REL * 1e4, k


def test_rz1_on_axis_saddle(solved):
    from ion_gym_amd.physics import fields_rz as FR
    spec = _deck(-100.0)
    dc, _, _ = FR.spec_terms_rz(spec)
    b = solved["bases"]
    A = sum(dc[k] * b[k] / 1e4 for k in b)
    h = spec.geometry.mm_per_gu
    ent0, ent1, l0, l1, ex0, ex1 = _edges(spec)
    cz = 0.5 * (l0 + l1)
    v_lens = A[int(round(cz / h)), 0]
    v_ent = A[int(round(5.0 / h)), 0]
    v_exit = A[int(round((ex1 - 2.0) / h)), 0]
    print(f"RZ-1 on axis: {v_ent:.2f} V (z=5) -> {v_lens:.2f} V (lens "
          f"centre) -> {v_exit:.2f} V (exit)")
    assert -100.0 < v_lens < -60.0
    assert abs(v_ent) < 5.0 and abs(v_exit) < 5.0


def test_rz5_three_tubes_shield_the_axis(solved):
    spec = _deck(-100.0)
    lab = solved["lab"]
    h = spec.geometry.mm_per_gu
    metal_z = (lab > 0).any(axis=1)
    starts = np.flatnonzero(np.diff(np.r_[0, metal_z.astype(int)]) == 1)
    ends = np.flatnonzero(np.diff(np.r_[metal_z.astype(int), 0]) == -1)
    assert len(starts) == len(ends) == 3
    tube_len = (ends[0] + 1 - starts[0]) * h
    bore = min(e.shapes[0]["r0_mm"] for e in spec.geometry.electrodes)
    ro = max(e.shapes[0]["r1_mm"] for e in spec.geometry.electrodes)
    assert tube_len > ro - bore
    # the lens tube alone at -100 V (the outer tubes grounded) holds the
    # axis at its centre below -60 V
    _, l0, l1 = _edges(spec)[1:4]
    cz = 0.5 * (l0 + l1)
    assert (solved["bases"][2] / 1e4 * -100.0)[int(round(cz / h)), 0] < -60.0


def test_rz4_retune_reuses_cached_bases(solved, monkeypatch):
    from ion_gym_amd.physics import solver2d as S
    from ion_gym_amd.physics.flight_conditions import (
        set_flight_conditions, solved_key)

    def _no_solve(*a, **k):
        raise AssertionError("a voltage change must not solve")
    monkeypatch.setattr(S, "laplace_factor", _no_solve)
    monkeypatch.setattr(S, "solve_rz", _no_solve)
    spec = _deck(-100.0)
    sk = solved_key(spec)
    t0 = time.perf_counter()
    set_flight_conditions(spec, voltages={"lens": -137.0}, solved_key=sk,
                          show=False)
    model = _model(solved, -137.0)
    dt = time.perf_counter() - t0
    print(f"RZ-4 retune to -137 V from the cached bases: {dt:.3f} s")
    assert dt < 0.5
    ref = _model(solved, -100.0)
    # the static field is linear in the lens volts
    assert np.allclose(model.EzA, ref.EzA * 1.37, rtol=1e-12,
                       atol=1e-12 * float(np.abs(ref.EzA).max()))


@pytest.mark.parametrize("precision", [None, "f32"])
def test_rz2_focus_then_over_focus(solved, precision):
    radii = []
    for v in (0.0, -80.0, -200.0, -400.0):
        got, cols, model, spec = _fly(solved, v, precision)
        radii.append(_beam_r(got, 36.0))
        if precision is None:
            o = _oracle(model, spec, solved["births"])
            _parity(got, o, cols)
    print(f"RZ-2 r(z=36 mm) at 0, -80, -200, -400 V "
          f"({precision or 'f64'}): {[round(r, 4) for r in radii]}")
    assert all(np.isfinite(radii))
    assert radii[2] < 0.4 * radii[0]
    assert radii[3] > radii[2]


def _parity(got, o, cols):
    assert len(got) == len(o["x"])
    for i, (tr, s) in enumerate(got):
        nr = int(o["nrec"][i])
        assert tr.shape == (nr, len(cols)), (i, tr.shape, nr)
        ref = o["rec"][i, :nr]
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-12)
        err = np.abs(tr - ref) / scale
        assert np.all(err <= REL), (i, cols[int(np.argmax(err.max(axis=0)))],
                                    float(err.max()))
        assert s["kind"] == int(o["kind"][i])
        assert abs(s["x_end"] - o["x"][i]) <= REL * 52.0


@pytest.mark.parametrize("precision", [None, "f32"])
def test_rz6_ions_stop_at_the_detector_focused(solved, precision):
    got, cols, model, spec = _fly(solved, -120.0, precision)
    births = solved["births"]
    x_end = np.array([s["x_end"] for _, s in got])
    zlen = spec.geometry.width_mm
    r_in = float(np.mean(np.hypot(births[:, 1], births[:, 2])))
    r_out = float(np.mean([math.hypot(s["y_end"], s["z_end"])
                           for _, s in got]))
    print(f"RZ-6 ({precision or 'f64'}): ions stop at x = "
          f"{x_end.min():.3f}-{x_end.max():.3f} mm, beam {r_in:.3f} -> "
          f"{r_out:.4f} mm")
    assert x_end.max() <= zlen and x_end.min() > zlen - 3.0
    assert r_out < 0.3 * r_in
    if precision is None:
        _parity(got, _oracle(model, spec, births), cols)
