"""Engine-backed planar fly function: the drop-in for make_planar_fly_fn
(ion_gym/physics/build_planar.py:1430-1685) over _fly_planar.

``make_planar_fly_fn(pack, births, spec)`` returns ``fly_fn(i) -> (traj,
summary)`` with the planar route's columns (BASE_CHANNELS, then the
requested channels in the route's own order), charge and model refusals,
birth refusals, the bounds and station planes in the kernel frame (an
anchored model's node i sits at i h + anchor), the SDS collision-count
rule, the drift extension after a box exit with its inset edge guard, and
the summary with r_end measured from the mean birth point.
``fly_fn.batch(indices)`` flies many ions in one launch.
"""
import math

import numpy as np

from ..engine import BASE_CHANNELS, PLANAR_CHANNELS, Flight
from .build_rz import RECORD_BUDGET_BYTES, shared_engine
from .drift_extension import edge_field_max, extend_ballistic
from .flight_inputs import DeckGas, ion_arrays
from .ion_envelope import (birth_splat, make_summary, mz_of, run_seed_base,
                           planar_birth_refusal)
from .stations import station_planes


def planar_edge_fields(pack):
    """Worst-case residual |E| (V/mm) on the four faces of a planar pack,
    each read one row in (build_planar.py:1548-1575): the static set
    exactly, each drive channel at |w| <= 1, or max |table| for tables."""
    tv = np.asarray(pack.get("tab_v", np.zeros(0)), float)
    tmax = float(np.max(np.abs(tv))) if tv.size else 1.0
    w = [(tmax if int(k) in (3, 4) else 1.0) for k in pack["ch_kind"]]
    ExK, EyK = np.asarray(pack["ExK"]), np.asarray(pack["EyK"])
    cx = [(pack["ExA"], 1.0)] + [(ExK[k], w[k]) for k in range(len(w))]
    cy = [(pack["EyA"], 1.0)] + [(EyK[k], w[k]) for k in range(len(w))]
    return {f"{ax}_{s}": 1e-3 * edge_field_max([cx, cy], i, s, inset=1)
            for i, ax in enumerate("xy") for s in ("lo", "hi")}


def make_planar_fly_fn(pack, births, spec, *, device=0, sds_stats=None,
                       anchor_mm=(0.0, 0.0), verbose=False):
    """(fly_fn, col_names) for a planar pack (ExA, EyA, ExK, EyK V/m, ele,
    mm, ch_*), births (N, 7) in the spec's frame and a FlightSpec."""
    bad = [c for c in spec.integration.record_channels
           if c not in PLANAR_CHANNELS]
    if bad:
        raise ValueError(f"planar route does not support record channel(s) "
                         f"{bad}; supported: {list(PLANAR_CHANNELS)}")
    if spec.transporter is not None:
        raise ValueError("the planar route does not implement the periodic "
                         "transporter")
    chans = [c for c in PLANAR_CHANNELS if c in spec.integration.record_channels]
    col_names = list(BASE_CHANNELS) + chans
    q_e = int(spec.source.charge)
    if q_e == 0:
        raise ValueError("source.charge = 0: an uncharged ion has no "
                         "electric acceleration; state the charge state.")
    gas = DeckGas(spec, sds_stats)
    sds_on = gas.sds_on
    dt = spec.dt_ns() * 1e-3
    rec_every = max(1, int(spec.integration.rec_every))
    n_steps = int(spec.integration.t_max_us / dt) + 2
    n_rows = n_steps // rec_every + 8
    anx, any_ = (float(v) for v in anchor_mm)
    # the user frame -> the kernel frame (node i at i h)
    flags, vals = spec.bounds.as_tuple()
    kvals = list(vals)
    kvals[0] -= anx
    kvals[1] -= anx
    kvals[2] -= any_
    kvals[3] -= any_
    # station planes in the kernel frame: the value along the plane's
    # axis and the window over its two transverse axes shifted by the
    # anchor (stations.py:94-130), z carrying none
    off = (anx, any_, 0.0)
    planes = station_planes(spec, off)
    for (a, v, sgn, win, kind) in (spec.planes or []):
        a = int(a)
        t1, t2 = [q for q in range(3) if q != a]
        w = (win[0] - off[t1], win[1] - off[t1], win[2] - off[t2],
             win[3] - off[t2])
        planes.append((a, float(v) - off[a], sgn, w, kind))
    flight = Flight(dt, spec.integration.t_max_us, rec_every=rec_every,
                    rec_rows=n_rows, channels=chans, gas=gas.kernel,
                    bounds=(tuple(flags), tuple(kvals)), planes=planes)
    births = np.asarray(births, np.float64)
    precision = spec.engine_precision()
    seed_base = run_seed_base(spec, verbose=verbose)
    ele = np.asarray(pack["ele"])
    h = float(pack["mm"])
    nx, ny = ele.shape
    axx, axy = 0.5 * (nx - 1) * h, 0.5 * (ny - 1) * h
    bx_mean, by_mean = float(births[:, 0].mean()), float(births[:, 1].mean())
    state = {}

    def _fields():
        if "fp" not in state:
            eng = shared_engine(device)
            state["eng"] = eng
            state["fp"] = eng.upload_planar(pack, precision)
        return state["eng"], state["fp"]

    def _finish(traj, kind, mz, ncol, i):
        if sds_on and "n_col" in col_names:
            traj[:, col_names.index("n_col")] = np.nan
        traj[:, 1] += anx
        traj[:, 2] += any_
        if kind == 1:
            xk, yk = float(traj[-1, 1]) - anx, float(traj[-1, 2]) - any_
            if xk < 0.0:
                fk, fn = "x_lo", "x low"
            elif xk / h > nx - 1:
                fk, fn = "x_hi", "x high"
            elif yk < 0.0:
                fk, fn = "y_lo", "y low"
            else:
                fk, fn = "y_hi", "y high"
            fa, fs = {"x_lo": (0, "lo"), "x_hi": (0, "hi"),
                      "y_lo": (1, "lo"), "y_hi": (1, "hi")}[fk]
            if "edge" not in state:
                state["edge"] = planar_edge_fields(pack)
            traj, kind, _ = extend_ballistic(
                traj, col_names, kind, spec,
                edge_e_vpermm=state["edge"][fk], face_axis=fa,
                face_side=fs, edge_face=fn,
                derived={"radius": lambda prev, row, ix, dt: math.sqrt(
                    (row[ix["x"]] - anx - axx) ** 2
                    + (row[ix["y"]] - any_ - axy) ** 2)},
                label=f"ion {i}")
        end = traj[-1]
        return traj, make_summary(
            kind=kind, tof=float(end[0]), mz=mz, x_end=float(end[1]),
            y_end=float(end[2]), z_end=float(end[3]),
            r_end=float(math.hypot(end[1] - bx_mean, end[2] - by_mean)),
            n_col=None if sds_on else ncol)

    def batch(indices):
        """[(traj, summary)] for the given ion indices, flown together."""
        idx = np.asarray(list(indices), np.int64)
        out = [None] * len(idx)
        fly_idx = []
        for k, i in enumerate(idx):
            b = births[i]
            err = planar_birth_refusal(ele, h, b[0] - anx, b[1] - any_, i)
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
        per_ion_bytes = n_rows * len(col_names) * 8
        chunk = max(1, min(len(fly_idx), RECORD_BUDGET_BYTES // per_ion_bytes))
        for c0 in range(0, len(fly_idx), chunk):
            ks = fly_idx[c0:c0 + chunk]
            ions, env = ion_arrays(spec, gas, births, idx[ks], seed_base,
                                   offset=(anx, any_))
            res = eng.fly_batch(fp, flight, ions)
            if np.any(res["fate"] == -3):
                j = int(np.nonzero(res["fate"] == -3)[0][0])
                raise ValueError(f"ion {int(idx[ks][j])}: hard-sphere "
                                 f"collision: no gas partner accepted in "
                                 f"MAX_PARTNER_TRIES draws")
            for j, k in enumerate(ks):
                nrec = int(res["nrec"][j])
                traj = res["records"][j, :nrec].copy()
                out[k] = _finish(traj, int(res["fate"][j]), env[j].mz,
                                 int(res["ncol"][j]), int(idx[k]))
        return out

    def fly_fn(i):
        return batch([i])[0]

    fly_fn.batch = batch
    fly_fn.col_names = col_names
    return fly_fn, col_names
