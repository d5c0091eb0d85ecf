"""A table (gate) drive on the r-z route holds its electrodes at
dc + amplitude_v * w(t) + offset_v before the step as well as after it:
ion_gym's regression test for the r-z gate (tests/test_rz_table_drive.py:
1-9, 55-80), whose r-z channel once dropped the table's first value and
the group offset.  The same deck: a gated plate (r 5-6 mm, dc 0, hold
table [10, 30] V at [0, 0.5] us, amplitude 1, offset 5) over a grounded
one (r 0-0.5 mm) on a 10 x 6 mm r-z domain at 0.25 mm; an ion of m/z 2000
at rest at (5, 3) mm.

CPU: the composed static potential is 15 V (= 0 + 1 x 10 + 5) on the
gated nodes and the gate set adds 20 V there from 0.5 us (the
reference's first assertion), and the oracle's ion has been pushed off
its radius before the gate (vy != 0 at 0.4 us, its second).  GPU: the
engine-backed fly function's ion likewise, its records equal to the
oracle's at 1e-9.
"""
import numpy as np
import pytest

from oracle import oracle as O

DC, AMP, V0, V1, OFFSET, TAU_US = 0.0, 1.0, 10.0, 30.0, 5.0, 0.5


def _deck():
    from ion_gym_amd.physics.flight_spec import (
        CollisionSpec, ElectrodeSpec, FlightSpec, GeometrySpec,
        IntegrationSpec, RFGroupSpec, SourceSpec)
    gate = RFGroupSpec(name="gt", amplitude_v=AMP, frequency_hz=0.0,
                       waveform="table", table_t_us=[0.0, TAU_US],
                       table_v=[V0, V1], interp="hold", offset_v=OFFSET)
    geo = GeometrySpec(
        width_mm=10.0, height_mm=6.0, mm_per_gu=0.25, symmetry="cylindrical",
        rf_groups=[gate],
        electrodes=[ElectrodeSpec(name="gated", dc=DC, rf_groups=["gt"],
                                  shapes=[{"type": "rect", "z0_mm": 0.0,
                                           "z1_mm": 10.0, "r0_mm": 5.0,
                                           "r1_mm": 6.0}]),
                    ElectrodeSpec(name="ground", dc=0.0,
                                  shapes=[{"type": "rect", "z0_mm": 0.0,
                                           "z1_mm": 10.0, "r0_mm": 0.0,
                                           "r1_mm": 0.5}])])
    src = SourceSpec(distribution="point", n_ions=1, x0_mm=5.0, y0_mm=3.0,
                     direction=[1.0, 0.0, 0.0], ke_lo=0.0, ke_hi=0.0,
                     mz_list=[2000.0], tob_span_us=0.0, seed=1)
    return FlightSpec(source=src, collisions=CollisionSpec(enabled=False),
                      integration=IntegrationSpec(dt_ns=1.0, t_max_us=1.0,
                                                  rec_every=1),
                      geometry=geo, name="gate_rz")


@pytest.fixture(scope="module")
def model():
    from ion_gym_amd import configs as C
    from ion_gym_amd.physics import fields_rz as FR
    from ion_gym_amd.physics.solver2d import solve_rz
    spec = _deck()
    lab = C.rz_deck_labels(spec)
    keys = sorted(int(v) for v in np.unique(lab) if v)
    bases = dict(zip(keys, solve_rz(lab, [{k: 1e4} for k in keys],
                                    spec.geometry.mm_per_gu)))
    return spec, lab, bases, FR.compose_spec_rz(spec, bases,
                                                lab.astype(np.int32))


def _births():
    from oracle import births as OB
    b = OB.generate_births(_deck().source, 1)
    assert np.array_equal(b, [[5.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0]])
    return b


def _vy_at(traj, t_us):
    row = np.searchsorted(traj[:, 0], t_us)
    return traj[row, 5]


def test_pre_gate_potential_and_the_gate_step(model):
    from ion_gym_amd.physics import fields_rz as FR
    spec, lab, bases, m = model
    dc, members, groups = FR.spec_terms_rz(spec)
    keys = sorted(bases)
    tA, tK, tG, drives, tau = FR.compose_terms_rz(keys, dc, members, groups)
    fas = [bases[k] / 1e4 for k in keys]
    A = FR.apply_terms(fas, tA, lab.shape)
    G = FR.apply_terms(fas, tG, lab.shape)
    gated = lab == 1
    assert gated.any() and tau == TAU_US and drives == []
    assert np.allclose(A[gated], DC + AMP * V0 + OFFSET, rtol=0, atol=1e-9)
    assert np.allclose(G[gated], AMP * (V1 - V0), rtol=0, atol=1e-9)
    assert m.tau_gate == TAU_US and m.EzG is not None


def test_oracle_ion_feels_the_pre_gate_field(model):
    from ion_gym_amd.physics.constants import E_CHG, KG_AMU
    spec, lab, bases, m = model
    b = _births()
    mass = np.array([2000.0])
    ions = dict(x=b[:, 0], y=b[:, 1], z=b[:, 2], vx=b[:, 3], vy=b[:, 4],
                vz=b[:, 5], tob=b[:, 6], m_amu=mass,
                acc=E_CHG / (mass * KG_AMU), sigma_m2=np.zeros(1),
                seed=np.array([1], np.int64))
    o = O.fly_rz(m.pack(), dict(dt=1e-3, t_max=1.0, rec_every=1,
                                rec_rows=1100, flags=[1, 1, 0, 0, 0, 1] +
                                [0] * 7), ions)
    tr = o["rec"][0, :int(o["nrec"][0])]
    # the plate above (15 V) pushes the positive ion toward the axis
    assert _vy_at(tr, 0.4) < 0.0


@pytest.mark.gpu
def test_engine_ion_feels_the_pre_gate_field(model):
    from ion_gym_amd.physics.build_rz import make_rz_fly_fn
    from ion_gym_amd.physics.constants import E_CHG, KG_AMU
    spec, lab, bases, m = model
    b = _births()
    print("r-z gate table, one ion, FP64", flush=True)
    fly_fn, cols = make_rz_fly_fn(m, b, spec)
    tr, s = fly_fn(0)
    assert _vy_at(tr, 0.4) < 0.0
    mass = np.array([2000.0])
    ions = dict(x=b[:, 0], y=b[:, 1], z=b[:, 2], vx=b[:, 3], vy=b[:, 4],
                vz=b[:, 5], tob=b[:, 6], m_amu=mass,
                acc=E_CHG / (mass * KG_AMU), sigma_m2=np.zeros(1),
                seed=np.array([1], np.int64))
    o = O.fly_rz(m.pack(), dict(dt=1e-3, t_max=1.0, rec_every=1,
                                rec_rows=int(1.0 / 1e-3) + 8,
                                flags=[1, 1, 0, 0, 0, 1] + [0] * 7), ions)
    ref = o["rec"][0, :int(o["nrec"][0])]
    assert tr.shape == ref.shape
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-12)
    assert float((np.abs(tr - ref) / scale).max()) <= 1e-9
    assert s["kind"] == int(o["kind"][0])
