"""Engine-backed r-z fly function: the drop-in for make_rz_fly_fn.

ion_gym/physics/build_rz.py:561-777 returns ``fly_fn(i) -> (traj,
summary)`` that flies ion i through the numba kernel _fly_rec_full.  This
module returns a function with the same signature and results, flown by the
MI355X engine, plus ``fly_fn.batch(indices)`` that flies many ions in one
launch; ion_gym_amd.physics.ensemble_driver takes that path when it is
present (the reference's sim_build propagates fly-function attributes,
sim_build.py:834-842).
"""
import math
import threading

import numpy as np

from ..engine import BASE_CHANNELS, RZ_CHANNELS, Engine, Flight
from .drift_extension import rz_edge_fields, rz_extend
from .flight_inputs import DeckGas, ion_arrays
from .ion_envelope import (birth_splat, make_summary, mz_of,
                           run_seed_base, rz_birth_refusal)
from .stations import all_planes

_ENGINES = {}
_ENGINES_LOCK = threading.Lock()


def shared_engine(device=0):
    """One engine per device and process, created on first use."""
    with _ENGINES_LOCK:
        e = _ENGINES.get(device)
        if e is None:
            e = _ENGINES[device] = Engine(device)
        return e


# bytes of record buffer one launch may fill before the batch is split
RECORD_BUDGET_BYTES = 1 << 30


def make_rz_fly_fn(model, births, spec, *, device=0, sds_stats=None,
                   verbose=False):
    """(fly_fn, col_names) for an RZModel, births (N, 7) and a FlightSpec.
    Refusals match the reference: an unknown record channel, charge 0, an
    unknown collision or sigma model.  The step is the deck's (or the one
    derived from its drives), the field storage its channel_dtype (or the
    spec's engine precision), the cross section of each ion its sigma
    model's."""
    chans = [c for c in RZ_CHANNELS if c in spec.integration.record_channels]
    bad = [c for c in spec.integration.record_channels
           if c not in RZ_CHANNELS]
    if bad:
        raise ValueError(f"r-z tracer cannot record channel(s) {bad}; the "
                         f"r-z kernel produces {list(RZ_CHANNELS)}")
    col_names = BASE_CHANNELS + chans
    q_e = int(spec.source.charge)
    if q_e == 0:
        raise ValueError("source.charge = 0: an uncharged ion has no "
                         "electric acceleration; state the charge state.")
    gas = DeckGas(spec, sds_stats)
    sds_on = gas.sds_on
    dt = spec.dt_ns() * 1e-3
    t_max = spec.integration.t_max_us
    rec_every = max(int(spec.integration.rec_every), 1)
    n_rows = int(t_max / dt / rec_every) + 8
    n_rows = min(max(n_rows, 16), 4_000_000)
    flight = Flight(dt, t_max, rec_every=rec_every, rec_rows=n_rows,
                    channels=chans, gas=gas.kernel,
                    bounds=spec.bounds.as_tuple(), planes=all_planes(spec))
    births = np.asarray(births, np.float64)
    precision = spec.engine_precision()
    # the run's seed policy, resolved once per build (build_rz.py:670-673)
    seed_base = run_seed_base(spec, verbose=verbose)
    state = {}
    # residual |E| on the open faces, the drift extension's guard input
    # (measured once per model, build_rz.py:650-669)
    edge_fields = rz_edge_fields(model)

    def _fields():
        if "fp" not in state:
            eng = shared_engine(device)
            state["eng"] = eng
            state["fp"] = eng.upload_rz(model.pack(), precision)
        return state["eng"], state["fp"]

    def batch(indices):
        """[(traj, summary)] for the given ion indices, flown together."""
        idx = np.asarray(list(indices), np.int64)
        out = [None] * len(idx)
        fly_idx = []
        for k, i in enumerate(idx):
            b = births[i]
            r = math.hypot(b[1], b[2])
            err = rz_birth_refusal(model.ele, model.mm_per_gu, b[0], r, i)
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
            ions, env = ion_arrays(spec, gas, births, idx[ks], seed_base)
            res = eng.fly_batch(fp, flight, ions)
            if np.any(res["fate"] == -3):
                j = int(np.nonzero(res["fate"] == -3)[0][0])
                raise ValueError(f"ion {int(idx[ks][j])}: hard-sphere "
                                 f"collision: no gas partner accepted in "
                                 f"MAX_PARTNER_TRIES draws; the ion velocity "
                                 f"or the gas thermal speed is not finite")
            for j, k in enumerate(ks):
                nrec = int(res["nrec"][j])
                traj = res["records"][j, :nrec].copy()
                if sds_on and "n_col" in col_names:
                    traj[:, col_names.index("n_col")] = np.nan
                kind = int(res["fate"][j])
                # declared bounds/stations beyond an escape face: one exact
                # ballistic record (build_rz.py:746-770)
                traj, kind, _ = rz_extend(traj, col_names, kind, spec, model,
                                          edge_fields,
                                          label=f"ion {int(idx[k])}")
                end = traj[-1]
                out[k] = (traj, make_summary(
                    kind=kind, tof=float(end[0]),
                    mz=env[j].mz, x_end=float(end[1]), y_end=float(end[2]),
                    z_end=float(end[3]),
                    r_end=float(math.hypot(end[2], end[3])),
                    n_col=None if sds_on else int(res["ncol"][j])))
        return out

    def fly_fn(i):
        return batch([i])[0]

    fly_fn.batch = batch
    fly_fn.col_names = col_names
    return fly_fn, col_names
