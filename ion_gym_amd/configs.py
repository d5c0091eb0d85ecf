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


def drift_rz(n_ions=10000, seed=4000, coarse=1, e_td=10.0, mz=500.0,
             K0=1.5, P_torr=4.0, T_k=300.0, dt_frac=0.1):
    """Drift tube with SDS: a 100-ring stack (0.4 mm rings at 0.8 mm pitch,
    inner radius 8 mm) on an 801 x 101 r-z grid @ 0.1 mm whose linear
    ladder gives a uniform 1 V/mm base field; N2 at 4 Torr and 300 K; the
    ion's declared K0 = 1.5 cm^2/(V s); births on the axis region at
    x = 5 mm with 300 K velocities; end plates at 80 V and 0 V close the
    stack (the exit plate at x = 80 mm is the detector).  The field
    strength E/N (Td) sets the
    static set's scale (``set_scale``); dt = dt_frac / gamma."""
    from .physics import sds as sds_mod
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
    n_gas = P_torr * (101325.0 / 760.0) / (K_B * T_k)   # 1/m^3
    E_vpm = e_td * 1e-21 * n_gas                         # V/m
    scale = E_vpm * 1e-3 / 1.0                           # base is 1 V/mm
    g = gas_mod.gas_kernel_scalars("N2", T_k=T_k, P_torr=P_torr)
    stats = sds_mod.default_statistics()
    p = sds_mod.transport_params(mz, 1, g["m_gas"], T_k, P_torr, K0,
                                 stats=stats)
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
    gas = dict(model=2, T_k=g["T_k"], P_pa=g["P_pa"], c_star=g["c_star"],
               c_bar=g["c_bar"], sig1d=g["sig1d"], m_gas=g["m_gas"],
               flow=(0.0, 0.0, 0.0), sds_stats=stats)
    return dict(name="drift_rz", route="rz", model=model, ions=ions, dt=dt,
                t_max=2000.0, gas=gas, scale=scale, E_vpmm=E_vpm * 1e-3,
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
