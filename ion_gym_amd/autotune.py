"""Launch autotuner for the flight kernels.

The flight kernels run one ion per lane; what is left to choose per
workload is how the lanes are launched:

* ``block``: lanes per workgroup (64, 128, 192, 256).  Smaller blocks let
  the dispatcher fill the 256 CUs more evenly when ions retire unevenly.
* ``occupancy``: the register-budget variant of the flight kernels (at
  least 1, 2 or 3 waves per SIMD, and 4 for the planar FP32-basis vacuum
  step; a budget below the kernel's own need keeps the rest in scratch).
  The 3-D SDS kernel has one variant.
* ``coll_batch``: how many parked lanes a 64-lane wave gathers before it
  scatters their hard-sphere collisions together (FP64 and FP32-storage
  modes; the FP32-basis mode tries partners one per loop pass and ignores
  it).

None of them changes a result (tests/test_gpu_autotune.py flies the same
ions under every setting and compares bit for bit), so the tuner may pick
purely on time.  It flies a sample of the caller's own ions for a short
window under each candidate, from the same births each time, times the
launch with the engine's HIP events, and keeps the fastest.  Winners are
cached on disk keyed by device, route, precision, set count, gas model
and record mode, so a workload is measured once per machine.

Two things the kernels' design leaves out of the search: ions per lane is
always one (a lane's state lives in registers for a whole window; two
ions per lane would double the register budget the measurements show is
already the limit), and there is no LDS tile (the r-z field packs are
L2/MALL resident and the measured bound is vector issue, not the gather).
"""
import itertools
import json
import os
import time

import numpy as np

from . import _lib as L

BLOCKS = (64, 128, 192, 256)
OCCUPANCIES = (1, 2, 3)
COLL_BATCHES = (1, 4, 16, 32)


def default_cache_path():
    return os.environ.get(
        "ION_GYM_AMD_LAUNCH_CACHE",
        os.path.join(os.path.expanduser("~"), ".cache", "ion_gym_amd",
                     "launch.json"))


def candidates(precision, gas_model, route=None):
    """The launch settings worth measuring for a precision, gas and
    route."""
    hs = gas_model == L.GAS_HS
    occs = OCCUPANCIES
    if route == L.ROUTE_XYZ and gas_model == L.GAS_SDS:
        occs = (1,)
    if (route == L.ROUTE_PLANAR and precision == "f32"
            and gas_model == L.GAS_VACUUM):
        occs = OCCUPANCIES + (4,)
    cbs = COLL_BATCHES if hs else (0,)
    return [dict(block=b, occupancy=o, coll_batch=c)
            for b, o, c in itertools.product(BLOCKS, occs, cbs)]


def cache_key(device_name, cu, fields, flight):
    gas = int(flight.gas.get("model", L.GAS_VACUUM))
    rec = int(flight.rec_rows > 0)
    return (f"{device_name}|{cu}|route{fields.route}|{fields.precision}|"
            f"sets{getattr(fields, 'n_sets', 1 + fields.n_chan)}|gas{gas}|"
            f"rec{rec}")


def _load_cache(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(path, data):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _sample(ions, n):
    return {k: np.asarray(v)[:n] for k, v in ions.items()}


def measure(engine, fields, flight, ions, cand, *, window=200, reps=2,
            load=None):
    """Milliseconds of one ``window``-step launch over ``ions`` with the
    launch settings ``cand`` (best of ``reps``; each rep re-flies the same
    births after one untimed warm-up window)."""
    fl = flight.launch(**cand)
    fl.max_steps = int(window)
    n = len(ions["x"])
    ens = engine.ensemble(n)
    try:
        best = float("inf")
        for _ in range(reps):
            (load or (lambda e, i: e.load(i)))(ens, ions)
            engine.fly(fields, fl, ens)          # warm-up window
            engine.fly(fields, fl, ens)
            best = min(best, engine.last_fly_ms())
        return best
    finally:
        ens.release()


def autotune(engine, fields, flight, ions, *, sample=131072, window=200,
             reps=2, cache_path=None, force=False, load=None, log=None,
             device=None):
    """Pick block, occupancy and coll_batch for flying ``ions`` (a births
    dict) through ``fields`` with ``flight``.  Returns a dict with the
    winning settings, their time and the measured table; ``flight.launch(
    **{k: r[k] for k in ("block", "occupancy", "coll_batch")})`` applies
    them.  ``load(ens, ions)`` may replace the plain ``ens.load`` (e.g. to
    install a tune table first).  ``device`` = (name, compute units)
    overrides the device query that keys the cache."""
    if device is None:
        from .engine import device_info
        device = device_info(engine.device)
    name, cu = device
    key = cache_key(name, cu, fields, flight)
    path = cache_path or default_cache_path()
    data = _load_cache(path)
    if not force and key in data:
        r = dict(data[key])
        r["cached"] = True
        return r
    sub = _sample(ions, sample)
    gas = int(flight.gas.get("model", L.GAS_VACUUM))
    table = []
    for cand in candidates(fields.precision, gas, fields.route):
        t0 = time.perf_counter()
        ms = measure(engine, fields, flight, sub, cand, window=window,
                     reps=reps, load=load)
        table.append(dict(cand, ms=ms))
        if log:
            log(f"[autotune] {cand} {ms:.3f} ms "
                f"({time.perf_counter() - t0:.2f} s)")
    best = min(table, key=lambda r: r["ms"])
    r = dict(block=best["block"], occupancy=best["occupancy"],
             coll_batch=best["coll_batch"], ms=best["ms"],
             ions=len(sub["x"]), window=int(window), table=table)
    data[key] = r
    _store_cache(path, data)
    r = dict(r)
    r["cached"] = False
    return r


def tuned(flight, result):
    """``flight`` with the autotuner's winning launch settings."""
    return flight.launch(**{k: result[k]
                            for k in ("block", "occupancy", "coll_batch")})
