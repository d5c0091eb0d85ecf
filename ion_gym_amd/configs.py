"""Synthetic workloads: the five flight configurations the engine is
measured and checked on, built from geometry, a Laplace solve and seeded
births.

Each builder returns a dict with the field model/pack, the flight
parameters and the births (passed as data, the reference's
``distribution: "file"`` route, ion_gym/physics/sim_build.py:89-139).

  einzel_rz   einzel lens r-z, vacuum, FP64 (401 x 101 @ 0.1 mm)
  quad_3d     3-D quadrupole, RF + DC, vacuum (161 x 161 x 401 @ 0.05 mm)
  funnel_rz   ion funnel r-z with hard-sphere He (1201 x 201 @ 0.1 mm)
  drift_rz    drift tube r-z with SDS N2 (801 x 101), many ensembles
  slim_3d     3-D traveling-wave board, 11 field sets, HS N2
"""
import math

import numpy as np

from .physics import collision3d as gas_mod
from .physics.constants import E_CHG, K_B, KG_AMU
from .physics.fields_rz import DriveGroup, RZModel, build_field_aware
from .physics.solver2d import solve_rz


def _grid_labels(nz, nr, h, electrodes):
    """electrodes: list of (label, z0, z1, r0, r1) rectangles (mm) painted
    onto node centres."""
    lab = np.zeros((nz, nr), np.int32)
    z = np.arange(nz) * h
    r = np.arange(nr) * h
    eps = 1e-9
    for L, z0, z1, r0, r1 in electrodes:
        iz = (z >= z0 - eps) & (z <= z1 + eps)
        ir = (r >= r0 - eps) & (r <= r1 + eps)
        lab[np.ix_(iz, ir)] = L
    return lab


def _speed_mm_us(ke_ev, m_amu):
    return np.sqrt(2.0 * ke_ev * E_CHG / (m_amu * KG_AMU)) * 1e-3


def _model_from_potentials(labels, h, A, chans, ch_kind, ch_om, ch_ph,
                           ch_duty):
    nz, nr = labels.shape
    Z = np.arange(nz) * h
    U = np.arange(nr) * h
    _, Ue, EzA, EuA = build_field_aware(Z, U, A, labels)
    EzK, EuK = [], []
    for B in chans:
        _, _, a, b = build_field_aware(Z, U, B, labels)
        EzK.append(a)
        EuK.append(b)
    EzK = np.stack(EzK) if EzK else np.zeros((0,) + EzA.shape)
    EuK = np.stack(EuK) if EuK else np.zeros((0,) + EuA.shape)
    return RZModel(EzA=EzA, EuA=EuA, EzK=EzK, EuK=EuK,
                   ch_kind=np.asarray(ch_kind, np.int64),
                   ch_om=np.asarray(ch_om, np.float64),
                   ch_ph=np.asarray(ch_ph, np.float64),
                   ch_duty=np.asarray(ch_duty, np.float64), ele=labels,
                   u0=float(Ue[0]), mm_per_gu=h)


def einzel_rz(n_ions=10000, seed=0, coarse=1):
    """Einzel lens: three coaxial tubes (inner radius 5 mm, 1 mm wall,
    10 mm long, 2 mm gaps) at 0 / +600 / 0 V; m/z 100, z = +1; births on
    the plane x = 1 mm, r uniform in 0-2 mm, KE ~ N(100, 1) eV along +x;
    DC, so dt = 1 ns; vacuum; fly to fate.  ``coarse`` > 1 divides the
    grid resolution (tests)."""
    h = 0.1 * coarse
    nz, nr = 400 // coarse + 1, 100 // coarse + 1
    elec = [(1, 4.0, 14.0, 5.0, 6.0), (2, 16.0, 26.0, 5.0, 6.0),
            (3, 28.0, 38.0, 5.0, 6.0)]
    lab = _grid_labels(nz, nr, h, elec)
    (A,) = solve_rz(lab, [{1: 0.0, 2: 600.0, 3: 0.0}], h)
    model = _model_from_potentials(lab, h, A, [], [], [], [], [])
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, 2.0, n_ions)
    th = rng.uniform(0.0, 2 * math.pi, n_ions)
    ke = rng.normal(100.0, 1.0, n_ions)
    m = np.full(n_ions, 100.0)
    v = _speed_mm_us(ke, m)
    births = dict(x=np.full(n_ions, 1.0), y=r * np.cos(th), z=r * np.sin(th),
                  vx=v, vy=np.zeros(n_ions), vz=np.zeros(n_ions),
                  tob=np.zeros(n_ions))
    q = 1
    ions = dict(births, m_amu=m, acc=q * E_CHG / (m * KG_AMU),
                sigma_m2=np.zeros(n_ions),
                seed=seed + np.arange(n_ions, dtype=np.int64))
    return dict(name="einzel_rz", route="rz", model=model, ions=ions,
                dt=1.0 * 1e-3, t_max=50.0, gas=None, charge=q)


def funnel_rz(n_ions=1_000_000, seed=3, coarse=1, dt_ns=0.25):
    """Ion funnel: 100 rings 0.5 mm thick at 1 mm pitch, inner diameter
    tapering linearly 25 -> 2 mm (3 mm radial depth); RF 1.0 MHz, 100 V
    zero-to-peak in opposite phase on alternate rings; 10 V/cm axial DC
    ladder; He at 2 Torr and 300 K, hard-sphere cross section 200 A^2;
    m/z log-uniform in 200-2000, z = +1; births uniform in a 10 mm disk on
    the entrance plane x = 8 mm with 300 K Maxwellian velocities; dt =
    0.25 ns (period/4000, below tau_collision/20 = 0.31 ns)."""
    h = 0.1 * coarse
    nz, nr = 1200 // coarse + 1, 200 // coarse + 1
    elec = []
    for i in range(100):
        z0 = 10.0 + i * 1.0
        rin = 12.5 + (1.0 - 12.5) * i / 99.0
        elec.append((i + 1, z0, z0 + 0.5, rin, rin + 3.0))
    lab = _grid_labels(nz, nr, h, elec)
    dc = {i + 1: 100.0 - (elec[i][1] - 10.0) * 1.0 for i in range(100)}
    ampA = {i + 1: (100.0 if i % 2 == 0 else 0.0) for i in range(100)}
    ampB = {i + 1: (100.0 if i % 2 == 1 else 0.0) for i in range(100)}
    A, BA, BB = solve_rz(lab, [dc, ampA, ampB], h)
    om = 1.0e6 * 1e-6 * 2 * math.pi
    model = _model_from_potentials(lab, h, A, [BA, BB], [0, 0], [om, om],
                                   [0.0, math.pi], [0.5, 0.5])
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(math.log(200.0), math.log(2000.0), n_ions))
    rr = 5.0 * np.sqrt(rng.uniform(0.0, 1.0, n_ions))
    th = rng.uniform(0.0, 2 * math.pi, n_ions)
    T = 300.0
    s1 = np.sqrt(K_B * T / (m * KG_AMU)) * 1e-3
    v = rng.normal(0.0, 1.0, (3, n_ions)) * s1
    births = dict(x=np.full(n_ions, 8.0), y=rr * np.cos(th),
                  z=rr * np.sin(th), vx=v[0], vy=v[1], vz=v[2],
                  tob=np.zeros(n_ions))
    q = 1
    g = gas_mod.gas_kernel_scalars("He", T_k=T, P_torr=2.0)
    gas = dict(model=1, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
               c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
               flow=(0.0, 0.0, 0.0))
    ions = dict(births, m_amu=m, acc=q * E_CHG / (m * KG_AMU),
                sigma_m2=np.full(n_ions, 2.0e-18),
                seed=seed + np.arange(n_ions, dtype=np.int64))
    return dict(name="funnel_rz", route="rz", model=model, ions=ions,
                dt=dt_ns * 1e-3, t_max=2000.0, gas=gas, charge=q)


def drift_winner_rz(n_ions=500, seed=60601, coarse=1, dt_ns=0.25):
    """ion_gym's certified drift-tube example (notebooks/data/
    20_drift_tube_matrix/drift_winner_spec.json): 40 rings 0.5 mm wide at
    2.1 mm pitch from x = 2 mm, inner radius 10 mm, outer 15 mm, DC ladder
    16.38 V -> 0 V in 0.42 V steps (0.2 V/mm, 6.9 Td), no RF; C60 (m/z
    720, z = +1) in He at 1 Torr and 298 K, hard sphere with a fixed
    1.25e-18 m^2; births on a 0.3 mm disc at x = 6 mm, KE uniform
    0.1-1.9 eV along +x, tob uniform over 2 us, seed 60601; 1600 us of
    flight at dt 0.25 ns (the deck's certified step, run_winner.py:49),
    records every 1 us.  The deck's 86.4 x 15.5 mm domain at 0.125 mm is
    691.2 cells long; the grid here is 691 cells (86.375 mm), the nearest
    conforming extent, which ends 1.9 mm past the last ring.

    Returns dict(model, spec (FlightSpec), source (SourceSpec), h)."""
    from .physics.flight_spec import (CollisionSpec, FlightSpec,
                                      IntegrationSpec, SourceSpec)
    h = 0.125 * coarse
    nz, nr = 691 // coarse + 1, 124 // coarse + 1
    elec = [(k + 1, 2.0 + 2.1 * k, 2.5 + 2.1 * k, 10.0, 15.0)
            for k in range(40)]
    lab = _grid_labels(nz, nr, h, elec)
    dc = {k + 1: round(16.38 - 0.42 * k, 10) for k in range(40)}
    (A,) = solve_rz(lab, [dc], h)
    model = _model_from_potentials(lab, h, A, [], [], [], [], [])
    src = SourceSpec(n_ions=n_ions, mz_list=[720.0], charge=1, seed=seed,
                     distribution="disc", x0_mm=6.0, r_mm=0.3, axis="x",
                     direction=[1.0, 0.0, 0.0], ke_lo=0.1, ke_hi=1.9,
                     temperature_k=0.0, tob_span_us=2.0)
    spec = FlightSpec(
        source=src,
        collisions=CollisionSpec(gas="He", T_k=298.0,
                                 P_torr=0.9999972366148532, model="hs",
                                 sigma_model="fixed", sigma_m2=1.25e-18),
        integration=IntegrationSpec(dt_ns=dt_ns, t_max_us=1600.0,
                                    rec_every=max(1, int(round(1000 / dt_ns))),
                                    record_channels=["speed", "ke_ev"]))
    return dict(name="drift_winner_rz", model=model, spec=spec, source=src,
                h=h)


def funnel_electrodes(coarse=2):
    """The funnel's electrodes as the reference's per-electrode cache holds
    them (build_rz.py:377-386): each ring's potential solved at 1e4 V with
    the others grounded, the DC ladder (100 V down at 1 V/mm), and two RF
    drive groups (1 MHz, 100 V, 0 and 180 degrees) on alternate rings.
    Returns (labels, h, bases {idx: phi}, dc, members, groups) for
    fields_rz.compose_rz and the device composition."""
    h = 0.1 * coarse
    nz, nr = 1200 // coarse + 1, 200 // coarse + 1
    elec = []
    for i in range(100):
        z0 = 10.0 + i * 1.0
        rin = 12.5 + (1.0 - 12.5) * i / 99.0
        elec.append((i + 1, z0, z0 + 0.5, rin, rin + 3.0))
    lab = _grid_labels(nz, nr, h, elec)
    keys = [e[0] for e in elec]
    sols = solve_rz(lab, [{k: 1e4} for k in keys], h)
    bases = dict(zip(keys, sols))
    dc = {i + 1: 100.0 - (elec[i][1] - 10.0) * 1.0 for i in range(100)}
    members = {i + 1: ("rfA",) if i % 2 == 0 else ("rfB",)
               for i in range(100)}
    groups = [DriveGroup("rfA", "sin", 1.0e6, 100.0, 0.0),
              DriveGroup("rfB", "sin", 1.0e6, 100.0, 180.0)]
    return lab, h, bases, dc, members, groups


def funnel_deck(coarse=2):
    """The funnel's electrodes as a deck: 100 ring electrodes (rectangles
    in (z, r), mm), a 100 -> 1 V linear DC ladder over their rung numbers
    (1 V/mm, the same volts as funnel_electrodes) and the two RF groups on
    alternate rings, on a cylindrical 120 x 20 mm domain at 0.1 * coarse
    mm.  Its bases come from rz_deck_labels and a Laplace solve; its
    voltages from the deck (fields_rz.compose_spec_rz)."""
    from .physics.flight_spec import (DCGroupSpec, ElectrodeSpec, FlightSpec,
                                      GeometrySpec, RFGroupSpec)
    rings = []
    for i in range(100):
        z0 = 10.0 + i * 1.0
        rin = 12.5 + (1.0 - 12.5) * i / 99.0
        rings.append(ElectrodeSpec(
            name=f"ring{i + 1:03d}",
            shapes=[{"type": "rect", "z0_mm": z0, "z1_mm": z0 + 0.5,
                     "r0_mm": rin, "r1_mm": rin + 3.0}],
            rf_groups=["rfA" if i % 2 == 0 else "rfB"],
            dc_group="ladder", dc_index=i))
    geo = GeometrySpec(
        width_mm=120.0, height_mm=20.0, mm_per_gu=0.1 * coarse,
        symmetry="cylindrical", electrodes=rings,
        dc_groups=[DCGroupSpec(name="ladder", v_in=100.0, v_out=1.0)],
        rf_groups=[RFGroupSpec("rfA", 1.0e6, 100.0, 0.0),
                   RFGroupSpec("rfB", 1.0e6, 100.0, 180.0)])
    return FlightSpec(geometry=geo, name=f"funnel_deck_c{coarse}")


def rz_deck_labels(spec):
    """(nz, nr) int labels of an r-z deck whose electrodes are 'rect'
    shapes: electrode k (1-based position, or its ``basis``) painted on
    the node centres inside its rectangles."""
    g = spec.geometry
    h = g.mm_per_gu
    nz = int(round(g.width_mm / h)) + 1
    nr = int(round(g.height_mm / h)) + 1
    rects = []
    for pos, el in enumerate(g.electrodes, start=1):
        idx = el.basis if el.basis is not None else pos
        for sh in el.shapes:
            if sh.get("type") != "rect":
                raise ValueError(f"electrode {el.name!r}: shape type "
                                 f"{sh.get('type')!r} is not 'rect'")
            rects.append((idx, sh["z0_mm"], sh["z1_mm"], sh["r0_mm"],
                          sh["r1_mm"]))
    return _grid_labels(nz, nr, h, rects)


def drives_for(model):
    """DriveGroup view of a composed model's channels (for reports)."""
    out = []
    for k in range(len(model.ch_kind)):
        out.append(DriveGroup(name=f"ch{k}", frequency_hz=model.ch_om[k] /
                              (2 * math.pi) * 1e6,
                              phase_deg=math.degrees(model.ch_ph[k])))
    return out


def _grad3(phi, h):
    """E = -grad(phi) in V/mm (np.gradient: central inside, one-sided at
    the array edges)."""
    gx, gy, gz = np.gradient(phi, h, h, h)
    return -gx, -gy, -gz


def quad_3d(n_ions=100_000, seed=2, coarse=1, a=0.02, q_ref_mz=500.0,
            rf_v=200.0, f_hz=1.0e6, n_cycles=2000):
    """3-D quadrupole: four round rods (r0 = 4 mm, rod radius 4.6 mm,
    20 mm long) on a 161 x 161 x 401 grid @ 0.05 mm (the 8 x 8 mm bore
    region), entrance and exit plates.  Bases (unit volts): rod pair A,
    rod pair B as the ideal quadrupole potentials 1/2 +- (x^2 - y^2) /
    (2 r0^2), and the two plates as exponential end fields.  RF 1 MHz at
    200 V0-p on A and B in opposite phase; DC U on A at the Mathieu point
    (a, q(m/z 500)), U = a V / (2 q).  m/z uniform in 400-600, births in a
    0.5 mm disk at z = 0.5 mm, KE U(4.5, 5.5) eV in a 2 degree cone;
    dt = period / 250 = 4 ns; 2,000 RF cycles."""
    h = 0.05 * coarse
    n = 160 // coarse + 1
    nzz = 400 // coarse + 1
    r0 = 4.0
    xs = np.arange(n) * h - 4.0
    zs = np.arange(nzz) * h
    X, Y, Z = np.meshgrid(xs, xs, zs, indexing="ij")
    quad = (X * X - Y * Y) / (2.0 * r0 * r0)
    lam = 1.0
    entr = np.exp(-Z / lam)
    exitp = np.exp(-(zs[-1] - Z) / lam)
    phiA = 0.5 + quad
    phiB = 0.5 - quad
    # rods: centre (+-8.6, 0) pair A, (0, +-8.6) pair B, radius 4.6
    R = 4.6
    c = r0 + R
    ele = (((X - c) ** 2 + Y ** 2 <= R * R) | ((X + c) ** 2 + Y ** 2 <= R * R)
           | (X ** 2 + (Y - c) ** 2 <= R * R) | (X ** 2 + (Y + c) ** 2
                                                <= R * R))
    om = f_hz * 1e-6 * 2 * math.pi
    Om = f_hz * 2 * math.pi
    m_ref = q_ref_mz * KG_AMU
    q = 4.0 * E_CHG * rf_v / (m_ref * Om ** 2 * (r0 * 1e-3) ** 2)
    U = a * rf_v / (2.0 * q)
    # static: U on pair A, plates at 0 V here (their bases keep the
    # channels available; entrance +2 V, exit -2 V give a gentle push)
    A = U * phiA + 2.0 * entr - 2.0 * exitp
    EA = _grad3(A, h)
    EB_A = _grad3(phiA, h)
    EB_B = _grad3(phiB, h)
    pack = dict(EAx=EA[0], EAy=EA[1], EAz=EA[2],
                ExK=np.stack([EB_A[0], EB_B[0]]),
                EyK=np.stack([EB_A[1], EB_B[1]]),
                EzK=np.stack([EB_A[2], EB_B[2]]), ele=ele, h_mm=h,
                ch_kind=np.array([0, 0], np.int64),
                ch_om=np.array([om, om]), ch_ph=np.array([0.0, math.pi]),
                ch_amp=np.array([rf_v, rf_v]), ch_off=np.zeros(2),
                ch_duty=np.full(2, 0.5))
    rng = np.random.default_rng(seed)
    m = rng.uniform(400.0, 600.0, n_ions)
    rr = 0.5 * np.sqrt(rng.uniform(0, 1, n_ions))
    th = rng.uniform(0, 2 * math.pi, n_ions)
    ke = rng.uniform(4.5, 5.5, n_ions)
    cone = math.radians(2.0)
    ct = rng.uniform(math.cos(cone), 1.0, n_ions)
    st = np.sqrt(1 - ct * ct)
    ph = rng.uniform(0, 2 * math.pi, n_ions)
    v = _speed_mm_us(ke, m)
    ions = dict(x=4.0 + rr * np.cos(th), y=4.0 + rr * np.sin(th),
                z=np.full(n_ions, 0.5), vx=v * st * np.cos(ph),
                vy=v * st * np.sin(ph), vz=v * ct, tob=np.zeros(n_ions),
                m_amu=m, acc=E_CHG / (m * KG_AMU),
                sigma_m2=np.zeros(n_ions),
                seed=seed + np.arange(n_ions, dtype=np.int64))
    dt_ns = 1e9 / f_hz / 250.0
    return dict(name="quad_3d", route="xyz", pack=pack, ions=ions,
                dt=dt_ns * 1e-9, t_max=n_cycles / f_hz, gas=None, U=U, q=q,
                a=a)


_DRIFT_MODELS = {}


def _drift_model(coarse):
    """The drift tube's field model (cached per grid coarsening): a
    100-ring stack (0.4 mm rings at 0.8 mm pitch, inner radius 8 mm) on an
    801 x 101 r-z grid @ 0.1 mm whose linear ladder gives a uniform 1 V/mm
    base field, closed by end plates at 80 V and 0 V (labels 101, 102)."""
    if coarse in _DRIFT_MODELS:
        return _DRIFT_MODELS[coarse]
    h = 0.1 * coarse
    nz, nr = 800 // coarse + 1, 100 // coarse + 1
    elec = []
    for i in range(100):
        z0 = 0.2 + 0.8 * i
        elec.append((i + 1, z0, z0 + 0.4, 8.0, 10.0))
    # entrance and exit plates close the stack, so the 1 V/mm ladder field
    # is uniform from end to end; the exit plate is the detector
    elec.append((101, 0.0, 0.0, 0.0, 10.0))
    elec.append((102, 80.0, 80.0, 0.0, 10.0))
    lab = _grid_labels(nz, nr, h, elec)
    # 1 V/mm: V(z) = 80 - z
    dc = {i + 1: 80.0 - (elec[i][1] + 0.2) for i in range(100)}
    dc[101] = 80.0
    dc[102] = 0.0
    (A,) = solve_rz(lab, [dc], h)
    model = _model_from_potentials(lab, h, A, [], [], [], [], [])
    _DRIFT_MODELS[coarse] = model
    return model


def _drift_ions(n_ions, seed, mz, e_td, K0, P_torr, T_k, dt_frac, stats,
                m_gas):
    """Births on the axis region at x = 5 mm with thermal velocities and the
    ions' SDS numbers; the field strength E/N (Td) as the 1 V/mm base
    field's scale; dt = dt_frac / gamma."""
    from .physics import sds as sds_mod
    n_gas = P_torr * (101325.0 / 760.0) / (K_B * T_k)   # 1/m^3
    E_vpm = e_td * 1e-21 * n_gas                         # V/m
    scale = E_vpm * 1e-3 / 1.0                           # base is 1 V/mm
    p = sds_mod.transport_params(
        mz, 1, m_gas, 0.366, T_k, P_torr, gas_name="N2",
        species=sds_mod.declared_K0(mz, "N2", K0), stats=stats)
    dt = dt_frac / p["damping"]
    rng = np.random.default_rng(seed)
    m = np.full(n_ions, mz)
    rr = 0.5 * np.sqrt(rng.uniform(0, 1, n_ions))
    th = rng.uniform(0, 2 * math.pi, n_ions)
    s1 = np.sqrt(K_B * T_k / (m * KG_AMU)) * 1e-3
    v = rng.normal(0.0, 1.0, (3, n_ions)) * s1
    ions = dict(x=np.full(n_ions, 5.0), y=rr * np.cos(th), z=rr * np.sin(th),
                vx=v[0], vy=v[1], vz=v[2], tob=np.zeros(n_ions), m_amu=m,
                acc=E_CHG / (m * KG_AMU), sigma_m2=np.zeros(n_ions),
                seed=seed + np.arange(n_ions, dtype=np.int64),
                sds=np.tile([p["damping"], p["mfp_mm"], p["V_mm_us"],
                             p["log_mr_ratio"]], (n_ions, 1)))
    return ions, dt, scale, E_vpm, p


def planar_quad(n_ions=2000, seed=12, coarse=1, a=0.02, q_ref_mz=500.0,
                rf_v=200.0, f_hz=1.0e6, n_cycles=50, vz_mm_us=0.5):
    """Planar route: the transverse section of a linear quadrupole guide
    (r0 = 4 mm) on a 161 x 161 grid @ 0.05 mm spanning +-4 mm about the
    axis, rods of radius 4.6 mm as metal.  Bases (V/m, amplitude baked into
    the channel grids as the planar route stores them): rod pair A and B as
    the ideal potentials 1/2 +- (X^2 - Y^2) / (2 r0^2); RF 1 MHz at
    rf_v V0-p on A and B in opposite phase; DC U on A at the Mathieu point
    (a, q(m/z 500)) in the static set.  m/z uniform in 400-600; births in
    a 0.5 mm disk about the axis with 0.05 mm/us transverse spread and an
    axial drift vz; dt = period / 250; n_cycles RF cycles."""
    h = 0.05 * coarse
    n = 160 // coarse + 1
    r0 = 4.0
    xs = np.arange(n) * h
    X, Y = np.meshgrid(xs - 4.0, xs - 4.0, indexing="ij")
    quad = (X * X - Y * Y) / (2.0 * r0 * r0)

    def grad(phi):
        gx, gy = np.gradient(phi, h)
        return -gx * 1e3, -gy * 1e3           # V/mm -> V/m

    EAx_u, EAy_u = grad(0.5 + quad)
    EBx_u, EBy_u = grad(0.5 - quad)
    R = 4.6
    c = r0 + R
    ele = (((X - c) ** 2 + Y ** 2 <= R * R) | ((X + c) ** 2 + Y ** 2 <= R * R)
           | (X ** 2 + (Y - c) ** 2 <= R * R) | (X ** 2 + (Y + c) ** 2
                                                <= R * R))
    om = f_hz * 1e-6 * 2 * math.pi
    Om = f_hz * 2 * math.pi
    m_ref = q_ref_mz * KG_AMU
    q = 4.0 * E_CHG * rf_v / (m_ref * Om ** 2 * (r0 * 1e-3) ** 2)
    U = a * rf_v / (2.0 * q)
    pack = dict(ExA=U * EAx_u, EyA=U * EAy_u,
                ExK=np.stack([rf_v * EAx_u, rf_v * EBx_u]),
                EyK=np.stack([rf_v * EAy_u, rf_v * EBy_u]),
                ele=ele.astype(np.float64), mm=h,
                ch_kind=np.array([0, 0], np.int64),
                ch_om=np.array([om, om]), ch_ph=np.array([0.0, math.pi]),
                ch_duty=np.full(2, 0.5))
    rng = np.random.default_rng(seed)
    m = rng.uniform(400.0, 600.0, n_ions)
    rr = 0.5 * np.sqrt(rng.uniform(0, 1, n_ions))
    th = rng.uniform(0, 2 * math.pi, n_ions)
    ions = dict(x=4.0 + rr * np.cos(th), y=4.0 + rr * np.sin(th),
                z=np.zeros(n_ions), vx=rng.normal(0, 0.05, n_ions),
                vy=rng.normal(0, 0.05, n_ions),
                vz=np.full(n_ions, vz_mm_us), tob=np.zeros(n_ions),
                m_amu=m, acc=E_CHG / (m * KG_AMU),
                sigma_m2=np.full(n_ions, 2e-18),
                seed=seed + np.arange(n_ions, dtype=np.int64))
    dt_us = 1e6 / f_hz / 250.0
    return dict(name="planar_quad", route="planar", pack=pack, ions=ions,
                dt=dt_us, t_max=n_cycles * 1e6 / f_hz, U=U, q=q, a=a)


def drift_rz(n_ions=10000, seed=4000, coarse=1, e_td=10.0, mz=500.0,
             K0=1.5, P_torr=4.0, T_k=300.0, dt_frac=0.1):
    """Drift tube with SDS: the 100-ring stack of _drift_model (uniform
    1 V/mm base field); N2 at 4 Torr and 300 K; the ion's declared
    K0 = 1.5 cm^2/(V s); births on the axis region at x = 5 mm with 300 K
    velocities; the exit plate at x = 80 mm is the detector.  The field
    strength E/N (Td) sets the static set's scale (``set_scale``);
    dt = dt_frac / gamma."""
    from .physics import sds as sds_mod
    model = _drift_model(coarse)
    g = gas_mod.gas_kernel_scalars("N2", T_k=T_k, P_torr=P_torr)
    stats = sds_mod.default_statistics()
    ions, dt, scale, E_vpm, p = _drift_ions(n_ions, seed, mz, e_td, K0,
                                            P_torr, T_k, dt_frac, stats,
                                            g["m_gas"])
    gas = dict(model=2, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
               c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
               flow=(0.0, 0.0, 0.0), sds_stats=stats)
    return dict(name="drift_rz", route="rz", model=model, ions=ions, dt=dt,
                t_max=2000.0, gas=gas, scale=scale, E_vpmm=E_vpm * 1e-3,
                K_mm2_per_V_us=p["K_mm2_per_V_us"], D=p["D_mm2_per_us"],
                transport=p)


def drift_batch_rz(n_ens=64, seed=4000, coarse=1, size_lo=1e3, size_hi=2e5,
                   td_lo=5.0, td_hi=20.0, mz_lo=100.0, mz_hi=1000.0, K0=1.5,
                   P_torr=4.0, T_k=300.0, dt_frac=0.1, t_max=2000.0):
    """Batched drift tube: n_ens ensembles for one launch on the drift_rz
    stack, each with its own size (log-uniform in size_lo-size_hi), field
    strength (uniform td_lo-td_hi Td, as the static set's scale) and m/z
    (uniform mz_lo-mz_hi), hence its own SDS numbers and dt = dt_frac /
    gamma; seeds seed + 1000 e.  Returns the model, the gas and, per
    ensemble, its births, tune (dt, t_max, set_scale) and the transport
    numbers (K, D) the drift and diffusion are checked against."""
    from .physics import sds as sds_mod
    model = _drift_model(coarse)
    g = gas_mod.gas_kernel_scalars("N2", T_k=T_k, P_torr=P_torr)
    stats = sds_mod.default_statistics()
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(size_lo), math.log(size_hi),
                                        n_ens))).astype(int)
    tds = rng.uniform(td_lo, td_hi, n_ens)
    mzs = rng.uniform(mz_lo, mz_hi, n_ens)
    ens = []
    for e in range(n_ens):
        ions, dt, scale, E_vpm, p = _drift_ions(
            int(sizes[e]), seed + 1000 * e, float(mzs[e]), float(tds[e]), K0,
            P_torr, T_k, dt_frac, stats, g["m_gas"])
        ens.append(dict(ions=ions, n=int(sizes[e]), td=float(tds[e]),
                        mz=float(mzs[e]),
                        tune=dict(dt=dt, t_max=t_max, set_scale=[scale]),
                        E_vpmm=E_vpm * 1e-3, K_mm2_per_V_us=p["K_mm2_per_V_us"],
                        D=p["D_mm2_per_us"]))
    gas = dict(model=2, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
               c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
               flow=(0.0, 0.0, 0.0), sds_stats=stats)
    return dict(name="drift_batch_rz", route="rz", model=model, gas=gas,
                ensembles=ens)


def drift_3d(n_ions=4000, seed=6000, coarse=1, e_vpmm=5.0, mz=500.0, K0=1.5,
             P_torr=4.0, T_k=300.0, dt_frac=0.1, rf_vpmm=0.0, rf_mhz=1.0,
             t_max_us=2000.0):
    """A 3-D SDS drift cell (the _fly3d_sds route): 41 x 41 x 301 nodes @
    0.1 mm (4 x 4 x 30 mm); a uniform static field e_vpmm along z (V/mm);
    optionally one sine drive channel of rf_vpmm along x at rf_mhz; a
    metal post (x, y in 2.6-3.4 mm, z in 10-12 mm) and a detector station
    plane at z = 25 mm (x, y in 0.5-3.5 mm); N2 at 4 Torr, 300 K; the ion's
    declared K0 = 1.5 cm^2/(V s); births near (2, 2, 1) mm with 300 K
    velocities and the ion's SDS numbers; dt = dt_frac / gamma in us.
    ``coarse`` > 1 divides the grid resolution (tests)."""
    from .physics import sds as sds_mod
    h = 0.1 * coarse
    nx = ny = 40 // coarse + 1
    nz = 300 // coarse + 1
    shape = (nx, ny, nz)
    EAx, EAy = np.zeros(shape), np.zeros(shape)
    EAz = np.full(shape, float(e_vpmm))
    X, Y, Z = np.meshgrid(np.arange(nx) * h, np.arange(ny) * h,
                          np.arange(nz) * h, indexing="ij")
    ele = ((np.abs(X - 3.0) <= 0.4) & (np.abs(Y - 3.0) <= 0.4)
           & (Z >= 10.0) & (Z <= 12.0))
    K = 1 if rf_vpmm else 0
    pack = dict(EAx=EAx, EAy=EAy, EAz=EAz, ele=ele, h_mm=h,
                ch_kind=np.zeros(K, np.int64),
                ch_om=np.full(K, 2 * math.pi * rf_mhz),
                ch_ph=np.zeros(K), ch_amp=np.ones(K), ch_off=np.zeros(K),
                ch_duty=np.full(K, 0.5))
    if K:
        pack.update(ExK=np.full((1,) + shape, float(rf_vpmm)),
                    EyK=np.zeros((1,) + shape), EzK=np.zeros((1,) + shape))
    else:
        pack.update(ExK=np.zeros((0,) + shape), EyK=np.zeros((0,) + shape),
                    EzK=np.zeros((0,) + shape))
    g = gas_mod.gas_kernel_scalars("N2", T_k=T_k, P_torr=P_torr)
    stats = sds_mod.default_statistics()
    p = sds_mod.transport_params(
        mz, 1, g["m_gas"], 0.366, T_k, P_torr, gas_name="N2",
        species=sds_mod.declared_K0(mz, "N2", K0), stats=stats)
    dt = dt_frac / p["damping"]
    rng = np.random.default_rng(seed)
    m = np.full(n_ions, mz)
    s1 = np.sqrt(K_B * T_k / (m * KG_AMU)) * 1e-3
    v = rng.normal(0.0, 1.0, (3, n_ions)) * s1
    ions = dict(x=2.0 + rng.uniform(-0.3, 0.3, n_ions),
                y=2.0 + rng.uniform(-0.3, 0.3, n_ions),
                z=np.full(n_ions, 1.0), vx=v[0], vy=v[1], vz=v[2],
                tob=rng.uniform(0.0, 0.5, n_ions), m_amu=m,
                acc=E_CHG / (m * KG_AMU), sigma_m2=np.zeros(n_ions),
                seed=seed + np.arange(n_ions, dtype=np.int64),
                sds=np.tile([p["damping"], p["mfp_mm"], p["V_mm_us"],
                             p["log_mr_ratio"]], (n_ions, 1)))
    gas = dict(model=2, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
               c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
               flow=(0.0, 0.0, 0.0), sds_stats=stats)
    planes = [(2, 25.0, 1.0, (0.5, 3.5, 0.5, 3.5), 6)]
    return dict(name="drift_3d", route="xyz", pack=pack, ions=ions, dt=dt,
                t_max=float(t_max_us), gas=gas, planes=planes,
                K_mm2_per_V_us=p["K_mm2_per_V_us"], D=p["D_mm2_per_us"],
                transport=p)


def hs_drift_rz(n_ions=2000, seed=5000, coarse=2, e_vpmm=1.0, mz=500.0,
                gas="He", P_torr=1.0, T_k=300.0, sigma_m2=2.0e-18,
                x_birth=76.0, dt_ns=None):
    """The drift tube of drift_rz with hard-sphere gas: uniform field
    e_vpmm V/mm (set_scale of the 1 V/mm ladder), births on the axis region
    at x_birth with thermal velocities, the exit plate at x = 80 mm as the
    detector (arrival = impact on it, fate 0).  dt = tau_collision / 20
    unless given."""
    base = drift_rz(n_ions=1, coarse=coarse)
    model = base["model"]
    g = gas_mod.gas_kernel_scalars(gas, T_k=T_k, P_torr=P_torr)
    tau_ns = gas_mod.collision_time_ns(gas, T_k, P_torr, sigma_m2)
    if dt_ns is None:
        dt_ns = tau_ns / gas_mod.HS_DT_PER_COLLISION_TIME
    rng = np.random.default_rng(seed)
    m = np.full(n_ions, mz)
    rr = 0.3 * np.sqrt(rng.uniform(0, 1, n_ions))
    th = rng.uniform(0, 2 * math.pi, n_ions)
    s1 = np.sqrt(K_B * T_k / (m * KG_AMU)) * 1e-3
    v = rng.normal(0.0, 1.0, (3, n_ions)) * s1
    ions = dict(x=np.full(n_ions, x_birth), y=rr * np.cos(th),
                z=rr * np.sin(th), vx=v[0], vy=v[1], vz=v[2],
                tob=np.zeros(n_ions), m_amu=m, acc=E_CHG / (m * KG_AMU),
                sigma_m2=np.full(n_ions, sigma_m2),
                seed=seed + np.arange(n_ions, dtype=np.int64))
    gas_d = dict(model=1, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
                 c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
                 flow=(0.0, 0.0, 0.0))
    return dict(name="hs_drift_rz", route="rz", model=model, ions=ions,
                dt=dt_ns * 1e-3, t_max=200.0, gas=gas_d, scale=e_vpmm,
                tau_ns=tau_ns)


# ----------------------------------------------------------------- SLIM
def _slab_fields(B, h, gap, ny, workers=None):
    """E = -grad(phi) of the Laplace solution between two parallel boards
    y = 0 and y = gap that both carry the board potential B(x, z)
    (periodic in x and z over B's padded extent): per Fourier mode k,
    phi_k(y) = B_k cosh(k (y - gap/2)) / cosh(k gap/2).  Returns Ex, Ey,
    Ez of shape (ny, Px, Pz) in V/mm per volt."""
    import scipy.fft as sfft
    Px, Pz = B.shape
    kx = 2 * np.pi * sfft.fftfreq(Px, d=h)[:, None]
    kz = 2 * np.pi * sfft.rfftfreq(Pz, d=h)[None, :]
    K = np.sqrt(kx * kx + kz * kz)
    Bh = sfft.rfft2(B, workers=workers)
    yc = np.arange(ny)[:, None, None] * h - 0.5 * gap
    ay = np.abs(yc)
    # cosh(K yc) / cosh(K gap/2) and sinh(K yc) / cosh(K gap/2) without
    # overflow
    e0 = np.exp(K[None] * (ay - 0.5 * gap)) / (1.0 + np.exp(-K[None] * gap))
    e2 = np.exp(-2.0 * K[None] * ay)
    rc = e0 * (1.0 + e2)
    rs = np.sign(yc) * e0 * (1.0 - e2)
    ph = Bh[None] * rc
    Ex = sfft.irfft2(-1j * kx[None] * ph, s=(Px, Pz), workers=workers)
    Ez = sfft.irfft2(-1j * kz[None] * ph, s=(Px, Pz), workers=workers)
    del ph
    Ey = sfft.irfft2(-(K[None] * Bh[None]) * rs, s=(Px, Pz), workers=workers)
    return Ex, Ey, Ez


def slim_3d(coarse=1, gap_mm=4.0, width_mm=10.0, length_mm=40.0, n_tw=48,
            n_phase=8, rf_v=100.0, rf_hz=1.0e6, dc_v=10.0, P_torr=4.0,
            T_k=300.0, workers=None):
    """3-D SLIM traveling-wave board: two parallel boards gap_mm apart
    (201 x 81 x 801 nodes @ 0.05 mm: x across the board, y across the gap,
    z along the path).  Each board carries, from the centre line out:
    48 TW electrodes in an 8-phase pattern across the central band
    x in [2.6, 7.4] mm (pitch length/48, 0.2 mm gaps), RF guard strips
    x in [1.6, 2.4] and [7.6, 8.4] (RF A) and [0.6, 1.4], [8.6, 9.4]
    (RF B, opposite phase), and DC guards x in [0, 0.4] and [9.6, 10].
    Field sets: the static set (the DC guards at dc_v) and K = 10 drive
    channels (8 TW phases: square, duty 0.5, unipolar amp*base + amp via
    ch_amp = ch_off = A/2; 2 RF: sin at rf_hz, 0 and pi), 11 in all.  The
    bases are the exact Laplace solution between the boards for the board
    potential (1 on a channel's electrodes, 0 on the others, linear across
    the 0.2 mm gaps), periodic over a grounded margin (FFT in x and z).
    N2 at P_torr, T_k with hard-sphere collisions; requests carry their own
    TW amplitude, speed, m/z and time step (slim_request)."""
    h = 0.05 * coarse
    nx = int(round(width_mm / h)) + 1
    ny = int(round(gap_mm / h)) + 1
    nz = int(round(length_mm / h)) + 1
    Px = 1 << int(math.ceil(math.log2(nx * 1.25)))
    Pz = 1 << int(math.ceil(math.log2(nz * 1.25)))
    ox = (Px - nx) // 2          # board origin inside the padded plane
    oz = (Pz - nz) // 2
    X = (np.arange(Px) - ox) * h
    Z = (np.arange(Pz) - oz) * h
    gap = 0.2
    pitch = length_mm / n_tw

    def band(lo, hi, v):
        return ((v >= lo) & (v <= hi)).astype(np.float64)

    zin = band(0.0, length_mm, Z)
    tw_x = band(2.6, 7.4, X)
    rfa_x = band(1.6, 2.4, X) + band(7.6, 8.4, X)
    rfb_x = band(0.6, 1.4, X) + band(8.6, 9.4, X)
    dc_x = band(-1e-9, 0.4, X) + band(9.6, width_mm + 1e-9, X)
    zz = Z - 0.0
    n_of_z = np.floor(zz / pitch).astype(int)
    in_el = np.abs(zz - (n_of_z + 0.5) * pitch) <= 0.5 * (pitch - gap)
    boards = []
    boards.append(dc_x[:, None] * zin[None, :])                # static
    for ph_i in range(n_phase):
        sel = (in_el & (n_of_z % n_phase == ph_i) & (zz >= 0)
               & (zz <= length_mm)).astype(np.float64)
        boards.append(tw_x[:, None] * sel[None, :])
    boards.append(rfa_x[:, None] * zin[None, :])
    boards.append(rfb_x[:, None] * zin[None, :])
    # linear ramps across the gaps: a box filter of the gap width in x and z
    w = max(1, int(round(gap / h)))
    kern = np.ones(w) / w

    def smooth(b):
        b = np.apply_along_axis(lambda r: np.convolve(r, kern, "same"), 0, b)
        return np.apply_along_axis(lambda r: np.convolve(r, kern, "same"),
                                   1, b)

    grids = []
    for b in boards:
        Ex, Ey, Ez = _slab_fields(smooth(b), h, gap_mm, ny, workers)
        crop = (slice(None), slice(ox, ox + nx), slice(oz, oz + nz))
        # (ny, nx, nz) -> (nx, ny, nz)
        grids.append([np.ascontiguousarray(E[crop].transpose(1, 0, 2))
                      for E in (Ex, Ey, Ez)])
        del Ex, Ey, Ez
    K = len(grids) - 1
    ele = np.zeros((nx, ny, nz), bool)
    # the two boards' electrode faces (a single node plane: the reference's
    # all-eight-corners rule never calls it metal, ions leave through the
    # box faces y = 0 and y = gap)
    face = (sum(boards[1:]) + boards[0])[ox:ox + nx, oz:oz + nz] > 0.5
    ele[:, 0, :] = face
    ele[:, -1, :] = face
    om_rf = rf_hz * 1e-6 * 2 * math.pi
    pack = dict(EAx=grids[0][0] * dc_v, EAy=grids[0][1] * dc_v,
                EAz=grids[0][2] * dc_v,
                ExK=np.stack([g[0] for g in grids[1:]]),
                EyK=np.stack([g[1] for g in grids[1:]]),
                EzK=np.stack([g[2] for g in grids[1:]]),
                ele=ele, h_mm=h,
                ch_kind=np.array([2] * n_phase + [0, 0], np.int64),
                ch_om=np.array([0.0] * n_phase + [om_rf, om_rf]),
                ch_ph=np.array([-2 * math.pi * q / n_phase
                                for q in range(n_phase)] + [0.0, math.pi]),
                ch_amp=np.array([10.0] * n_phase + [rf_v, rf_v]),
                ch_off=np.array([10.0] * n_phase + [0.0, 0.0]),
                ch_duty=np.full(K, 0.5))
    g = gas_mod.gas_kernel_scalars("N2", T_k=T_k, P_torr=P_torr)
    gas_d = dict(model=1, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
                 c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
                 flow=(0.0, 0.0, 0.0))
    return dict(name="slim_3d", route="xyz", pack=pack, gas=gas_d,
                pitch_mm=pitch, n_phase=n_phase, shape=(nx, ny, nz), h=h,
                P_torr=P_torr, T_k=T_k, width_mm=width_mm, gap_mm=gap_mm,
                length_mm=length_mm)


def slim_sigma_m2(mz):
    """Hard-sphere cross section of a singly charged ion of mass mz amu:
    2e-18 m^2 at 500 amu, scaling as m^(2/3)."""
    return 2.0e-18 * (np.asarray(mz, float) / 500.0) ** (2.0 / 3.0)


def slim_request(c, rng, n_ions=None, t_flight_us=2.0, rng_key=0):
    """One serving request on a slim_3d board: n ions (log-uniform in
    1e3-5e5 unless given) of one m/z (uniform 200-3000) born in a 1 mm box
    at the board centre 2 mm downstream of the entrance with 300 K
    velocities; the tune carries a TW amplitude A ~ U(10, 40) V (unipolar
    squares: amp = off = A/2), a TW speed v ~ U(50, 300) m/s (frequency
    v / (8 pitch)), the time step tau/20 for the ion's hard-sphere
    collision time and a flight time of t_flight_us.  Returns (ions,
    tune dict for serve.make_tune)."""
    n = int(n_ions) if n_ions is not None else int(
        round(math.exp(rng.uniform(math.log(1e3), math.log(5e5)))))
    mz = float(rng.uniform(200.0, 3000.0))
    A = float(rng.uniform(10.0, 40.0))
    v = float(rng.uniform(50.0, 300.0))
    f_per_us = v * 1e-3 / (c["n_phase"] * c["pitch_mm"])
    sig = float(slim_sigma_m2(mz))
    tau_ns = gas_mod.collision_time_ns("N2", c["T_k"], c["P_torr"], sig)
    dt_s = tau_ns / gas_mod.HS_DT_PER_COLLISION_TIME * 1e-9
    s1 = math.sqrt(K_B * c["T_k"] / (mz * KG_AMU)) * 1e-3
    vel = rng.normal(0.0, s1, (3, n))
    xc = 0.5 * c["width_mm"]
    yc = 0.5 * c["gap_mm"]
    ions = dict(x=xc + rng.uniform(-0.5, 0.5, n),
                y=yc + rng.uniform(-0.5, 0.5, n),
                z=2.0 + rng.uniform(-0.5, 0.5, n),
                vx=vel[0], vy=vel[1], vz=vel[2], tob=np.zeros(n),
                m_amu=np.full(n, mz), acc=np.full(n, E_CHG / (mz * KG_AMU)),
                sigma_m2=np.full(n, sig),
                seed=rng.integers(0, 2**40) + np.arange(n, dtype=np.int64))
    pk = c["pack"]
    K = len(pk["ch_kind"])
    nph = c["n_phase"]
    om = np.array(pk["ch_om"], float)
    om[:nph] = 2 * math.pi * f_per_us
    amp = np.array(pk["ch_amp"], float)
    off = np.array(pk["ch_off"], float)
    amp[:nph] = 0.5 * A
    off[:nph] = 0.5 * A
    tune = dict(dt=dt_s, t_max=t_flight_us * 1e-6, rng_key=int(rng_key),
                set_scale=np.ones(K + 1), ch_kind=np.array(pk["ch_kind"]),
                ch_om=om, ch_ph=np.array(pk["ch_ph"], float),
                ch_duty=np.array(pk["ch_duty"], float), ch_amp=amp,
                ch_off=off)
    return ions, tune
