"""Engine-backed 3-D fly function: the drop-in for the native 3-D route's
fly_fn (ion_gym/physics/build_stl3d.py:1181-1396) over tracer3d.fly3d and
fly3d_sds (tracer3d.py:630-785, 1029-1133).

``make_3d_fly_fn(pack, births, spec)`` returns ``fly_fn(i) -> (traj,
summary)`` with the reference's columns (BASE_CHANNELS, then the requested
optional channels in OPTIONAL_CHANNELS order, derived from the kernel's
recorded velocity and field as the reference derives them), summary keys
(kind, tof, mz, end point, ke_end, end velocity, r_end, n_col -- None under
SDS --, truncated, t_end_us), model dispatch (hard sphere or vacuum through
RK4, SDS through the SDS flight), the stations-then-bounds kernel planes
(stations.py:219-226), the transporter, the birth refusals, the drift
extension after a box exit, and the post-hoc bound scan for flights that
ended otherwise (build_stl3d.py:1102-1134).  ``fly_fn.batch(indices)`` flies
many ions in one launch.
"""
import math

import numpy as np

from ..engine import BASE_CHANNELS, Flight
from .build_rz import RECORD_BUDGET_BYTES, shared_engine
from .constants import E_CHG, KG_AMU
from .drift_extension import extend_ballistic, xyz_edge_fields, xyz_escape_face
from .flight_inputs import DeckGas, ion_arrays
from .ion_envelope import (birth_splat, make_summary, mz_of, run_seed_base,
                           xyz_birth_refusal)
from .stations import all_planes

# the optional channels the 3-D route fills, in the reference's
# OPTIONAL_CHANNELS order (io/sim_spec.py:119-138)
XYZ_CHANNELS = ("speed", "ke_ev", "ke_x", "ke_y", "ke_z", "e_field",
                "e_axial", "e_radial", "e_x", "e_y", "e_z", "radius", "n_col",
                "wrap_passes")
_IMPOSSIBLE = (math.inf, -math.inf, math.inf, -math.inf)


def kernel_planes(spec):
    """Stations, then the enabled bounds as whole-plane kills with fate 3
    (stations.py:199-226, tracer3d.pack_planes)."""
    out = all_planes(spec)
    flags, vals = spec.bounds.as_tuple()
    for k in range(6):
        if flags[k]:
            out.append((k // 2, float(vals[k]), 1.0 if k % 2 else -1.0,
                        _IMPOSSIBLE, 3))
    return out


def make_3d_fly_fn(pack, births, spec, *, device=0, sds_stats=None,
                   verbose=False):
    """(fly_fn, col_names) for a 3-D pack (EAx, EAy, EAz, ExK, EyK, EzK
    V/mm, ele, h_mm, ch_*), births (N, 7) and a FlightSpec."""
    opt = [c for c in XYZ_CHANNELS if c in spec.integration.record_channels]
    bad = [c for c in spec.integration.record_channels
           if c not in XYZ_CHANNELS]
    if bad:
        raise ValueError(f"3-D route cannot record channel(s) {bad}; it "
                         f"produces {list(XYZ_CHANNELS)}")
    col_names = list(BASE_CHANNELS) + opt
    q_e = int(spec.source.charge)
    if q_e == 0:
        raise ValueError("charge = 0: an uncharged ion has no electric "
                         "acceleration. State the charge state "
                         "(source.charge).")
    gas = DeckGas(spec, sds_stats)
    sds_on = gas.sds_on
    if sds_on and "wrap_passes" in opt:
        raise ValueError("record channel 'wrap_passes' was requested but "
                         "the collisions.model='sds' kernel does not "
                         "produce it")
    if sds_on and spec.transporter is not None:
        raise ValueError("the transporter belongs to the 3-D RK4 flight; "
                         "the SDS flight has none")
    integ = spec.integration
    dt_ns = spec.dt_ns()
    rec_every = max(int(integ.rec_every), 1)
    # rows a flight can write: the birth row, one per rec_every steps and
    # the terminal row (the reference's own estimate of the need,
    # build_stl3d.py:1165, plus one), at most max_records
    need = int(integ.t_max_us * 1e3 / (dt_ns * rec_every)) + 3
    n_rows = max(2, min(int(integ.max_records), need))
    # the flight clocks of the two 3-D kernels (tracer3d.py:765, 1117)
    if sds_on:
        dt, t_max = dt_ns * 1e-3, float(integ.t_max_us)
    else:
        dt, t_max = dt_ns * 1e-9, integ.t_max_us * 1e-6
    planes = kernel_planes(spec)
    flight = Flight(dt, t_max, rec_every=rec_every, rec_rows=n_rows,
                    gas=gas.kernel,
                    planes=planes, transporter=spec.transporter)
    births = np.asarray(births, np.float64)
    precision = spec.engine_precision()
    seed_base = run_seed_base(spec, verbose=verbose)
    metal = np.asarray(pack["ele"]).astype(bool)
    h = float(pack["h_mm"])
    shape = metal.shape
    flags, _ = spec.bounds.as_tuple()
    declares_downstream = bool(all_planes(spec)) or any(flags)
    state = {}

    def _fields():
        if "fp" not in state:
            eng = shared_engine(device)
            state["eng"] = eng
            state["fp"] = eng.upload_xyz(pack, precision)
        return state["eng"], state["fp"]

    def _traj(rec, nrec, mz, ncol):
        """The reference's trajectory columns from the kernel rows (t, x,
        y, z, vx, vy, vz, Ex, Ey, Ez[, passes]) (build_stl3d.py:1258-1318)."""
        r = rec[:nrec]
        tr = np.zeros((max(nrec, 1), len(col_names)))
        tr[:nrec, :7] = r[:, :7]
        vx, vy, vz = r[:, 4], r[:, 5], r[:, 6]
        ex, ey, ez = r[:, 7], r[:, 8], r[:, 9]
        sp = np.sqrt(vx * vx + vy * vy + vz * vz)
        m = mz * KG_AMU
        fill = {
            "speed": lambda: sp,
            "ke_ev": lambda: 0.5 * m * (sp * 1e3) ** 2 / E_CHG,
            "ke_x": lambda: 0.5 * m * (vx * 1e3) ** 2 / E_CHG,
            "ke_y": lambda: 0.5 * m * (vy * 1e3) ** 2 / E_CHG,
            "ke_z": lambda: 0.5 * m * (vz * 1e3) ** 2 / E_CHG,
            "e_field": lambda: np.sqrt(ex ** 2 + ey ** 2 + ez ** 2),
            "e_axial": lambda: ez,
            "e_radial": lambda: np.sqrt(ex ** 2 + ey ** 2),
            "e_x": lambda: ex, "e_y": lambda: ey, "e_z": lambda: ez,
            "radius": lambda: np.sqrt(r[:, 1] ** 2 + r[:, 2] ** 2),
            "n_col": lambda: (np.full(nrec, np.nan) if sds_on
                              else np.full(nrec, float(ncol))),
            "wrap_passes": lambda: (r[:, 10] if r.shape[1] > 10
                                    else np.zeros(nrec)),
        }
        for slot, name in enumerate(opt, start=7):
            tr[:nrec, slot] = fill[name]()
        return tr

    def _apply_bounds(tr, summ):
        """The post-hoc bound scan on the recorded rows for a flight the
        kernel did not end on a plane (build_stl3d.py:1102-1134)."""
        if int(summ.get("kind", -1)) == 3:
            return tr, summ
        bnd = spec.bounds
        cand = []
        for ax, c in (("x", 1), ("y", 2), ("z", 3)):
            if getattr(bnd, f"{ax}_min_on", False):
                cand.append((c, getattr(bnd, f"{ax}_min"), -1))
            if getattr(bnd, f"{ax}_max_on", False):
                cand.append((c, getattr(bnd, f"{ax}_max"), +1))
        if not cand or tr is None or not len(tr):
            return tr, summ
        hit = None
        for c, v, sgn in cand:
            idx = np.where((tr[:, c] - v) * sgn >= 0)[0]
            if idx.size and (hit is None or idx[0] < hit):
                hit = int(idx[0])
        if hit is not None:
            tr = tr[:max(hit, 1) + 1]
            summ = dict(summ, kind=3, tof=float(tr[-1, 0]),
                        x_end=float(tr[-1, 1]), y_end=float(tr[-1, 2]),
                        z_end=float(tr[-1, 3]),
                        r_end=float(math.hypot(tr[-1, 1], tr[-1, 2])))
        return tr, summ

    def batch(indices):
        """[(traj, summary)] for the given ion indices, flown together."""
        idx = np.asarray(list(indices), np.int64)
        out = [None] * len(idx)
        fly_idx = []
        for k, i in enumerate(idx):
            b = births[i]
            err = xyz_birth_refusal(metal, h, b[0], b[1], b[2], i)
            if err is not None:
                out[k] = birth_splat(err, col_names=col_names, birth=b,
                                     mz=mz_of(spec.source.mz_list,
                                              spec.source.n_ions, i),
                                     label=f"ion {i}")
            else:
                fly_idx.append(k)
        if not fly_idx:
            return out
        eng, fp = _fields()
        per_ion_bytes = n_rows * 11 * 8
        chunk = max(1, min(len(fly_idx), RECORD_BUDGET_BYTES // per_ion_bytes))
        for c0 in range(0, len(fly_idx), chunk):
            ks = fly_idx[c0:c0 + chunk]
            ions, env = ion_arrays(spec, gas, births, idx[ks], seed_base)
            res = eng.fly_batch(fp, flight, ions)
            if np.any(res["fate"] == -3):
                j = int(np.nonzero(res["fate"] == -3)[0][0])
                raise ValueError(f"ion {int(idx[ks][j])}: hard-sphere "
                                 f"collision: no gas partner accepted in "
                                 f"MAX_PARTNER_TRIES draws")
            for j, k in enumerate(ks):
                i = int(idx[k])
                nrec = int(res["nrec"][j])
                mz = env[j].mz
                ncol = int(res["ncol"][j])
                tr = _traj(res["records"][j], nrec, mz, ncol)
                kind = int(res["fate"][j])
                vx, vy, vz = (float(res[c][j]) for c in ("vx", "vy", "vz"))
                x_e, y_e, z_e = (float(res[c][j]) for c in ("x", "y", "z"))
                t_end = float(res["t_us"][j])
                tof = t_end
                ke = 0.5 * mz * KG_AMU * ((vx * 1e3) ** 2 + (vy * 1e3) ** 2
                                          + (vz * 1e3) ** 2) / E_CHG
                truncated = bool(nrec >= n_rows and nrec > 0
                                 and tr[nrec - 1, 0] < t_end - 1e-9)
                if (kind == 1 and nrec and declares_downstream
                        and not truncated):
                    fk, fa, fs, words = xyz_escape_face(shape, h,
                                                        (x_e, y_e, z_e))
                    if "edge" not in state:
                        state["edge"] = xyz_edge_fields(pack)
                    tr2, kind, xt = extend_ballistic(
                        tr[:nrec], col_names, kind, spec,
                        edge_e_vpermm=state["edge"][fk], face_axis=fa,
                        face_side=fs, edge_face=words,
                        derived={"radius": lambda prev, row, ix, dt:
                                 math.hypot(row[ix["x"]], row[ix["y"]])},
                        label=f"ion {i}")
                    if xt is not None:
                        tr = tr2
                        x_e, y_e, z_e = xt["end_mm"]
                        tof = float(tr[-1, 0])
                summ = make_summary(
                    kind=kind, tof=tof, mz=mz, x_end=x_e, y_end=y_e,
                    z_end=z_e, ke_end=ke, vx_end=vx, vy_end=vy, vz_end=vz,
                    r_end=float(math.hypot(x_e, y_e)),
                    n_col=None if sds_on else ncol, truncated=truncated,
                    t_end_us=t_end)
                out[k] = _apply_bounds(tr, summ)
        return out

    def fly_fn(i):
        return batch([i])[0]

    fly_fn.batch = batch
    fly_fn.col_names = col_names
    return fly_fn, col_names
