"""Throughput benchmark of the MI355X ion-flight engine.

Workload: the ion funnel r-z configuration with hard-sphere He
(1201 x 201 grid @ 0.1 mm, 100 rings, RF 1 MHz 100 V0-p on alternate
rings + 10 V/cm DC, He 2 Torr 300 K, sigma 200 A^2, 1,000,000 ions,
m/z log-uniform 200-2000, dt 0.25 ns), field bases in FP32 with FP64
positions and time (the engine's FP32-basis mode).  One "step" of the
benchmark is one window of S integrator steps for every live ion (one
ig_fly launch), with the ion state and the field resident in HBM.

Prints ONE JSON line.  ``value`` is ion-steps/s (live ion-steps, whole
job).  Beside it:

* the plain FP64 kernel's ion-steps/s on the same workload;
* the roofline of the flight kernel.  Its 16 MB field pack is served from
  L1/L2/MALL, not HBM, so the gather is measured against the gather rate
  the memory hierarchy reaches for the same access pattern: a probe kernel
  (csrc/probe_gather.hip) walks every ion's own sample-point sequence
  through an identical FP32 record array and does nothing else.  HBM
  traffic (FETCH_SIZE, WRITE_SIZE) and vector-issue counters
  (SQ_ACTIVE_INST_VALU, SQ_INSTS_VALU, SQ_WAVES) are read by two
  rocprofv3 counter runs of this script's own probe mode, each calibrated
  against a plain copy of known size in the same run; ``binding`` names
  the resource the kernel is closest to saturating;
* the CPU baseline (the oracle's C FP64 port of the reference kernels on
  the host cores, timed on a bounded sample, with the CPU model and the
  cost per ion-step per core beside the reference's own 87-101 ns);
* ``precision_funnel``: the FP32 basis against FP64 on this workload (the
  same births, one 4000-step window): distributions and means of the
  observables;
* the serving leg: flight requests on the 3-D SLIM traveling-wave board
  (201 x 81 x 801 @ 0.05 mm, 11 FP32 field sets, N2 4 Torr hard sphere,
  1e3-5e5 ions per request with their own TW amplitude, speed and m/z)
  arriving as Poisson streams at three offered loads (fractions of the
  engine's solo throughput), served by continuous batching; p50/p99
  request latency and the sustained ion-steps/s per load;
* ``precision``: the FP32-basis error against FP64 on the full-grid 3-D
  quadrupole (100,000 ions, 2,000 RF cycles).
"""
import argparse
import csv
import ctypes
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# the SDS legs fly ion_gym's own jump table (kept with the tests' fixtures)
os.environ.setdefault("ION_GYM_SDS_TABLE", os.path.join(
    HERE, "tests", "golden", "sds_jump_icdf.dat"))

HBM_PEAK_GBS = 8000.0
SIMDS = 1024                # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9            # MI355X peak engine clock
# the reference's own cost of an r-z ion-step with gas, one core
# (ion_gym/io/sim_spec.py:112-116)
REF_NS_PER_ION_STEP = (87.0, 101.0)


# Design choices behind the numbers (how batching, launch tuning and the
# roofline are done), stated on the benchmark line itself.
DESIGN_NOTES = {
    "batching": "one launch per window over every live ion of every "
                "ensemble (a per-ion tune row), consecutive windows "
                "replayed from a captured hipGraph; no persistent kernel: "
                "a lane already runs its ion's whole window in registers, "
                "so a persistent grid would add only a work queue",
    "autotuner": "searches block size, register budget (waves per SIMD) "
                 "and the hard-sphere collision batch; ions per lane stay "
                 "at one (the lane state fills the register budget) and "
                 "the r-z route has no LDS tile (its 16 MB pack is L2/MALL "
                 "resident and the bounds are the gather rate and vector "
                 "issue); the 3-D FP32 step keeps its cell's corner sums "
                 "in LDS",
    "roofline": "the r-z pack never leaves L2/MALL, so the kernel is "
                "priced against the measured gather rate of its own access "
                "pattern (probe_gather.hip, over the records as stored) "
                "with HBM traffic beside it",
    "channel_folding": "in the FP32-basis mode, drive channels of one "
                       "kind and frequency whose phases are whole half "
                       "periods apart are stored as one field set (the "
                       "grids combined in FP64, then rounded): the "
                       "funnel's opposite-phase RF pair flies as 2 sets "
                       "instead of 3, the SLIM board as 7 instead of 11; "
                       "the field is the unfolded one to the FP32 rounding "
                       "and alg_bytes_per_ion_step stays the reference "
                       "step's",
}


def gather_bytes_per_ion_step(n_sets, value_bytes):
    """Algorithmic gather bytes of one r-z ion-step of the reference's
    kick-drift-kick step (tracer_rz.py:207-440): two bilinear field
    evaluations of 2 components x n_sets sets x 4 corners, plus one
    metal-mask byte (E C c s + m).  The engine's kernel gathers half of
    it: the second evaluation of a step is the first of the next (same
    point, same lab time) and is kept in registers."""
    return 2 * (2 * n_sets) * 4 * value_bytes + 1


def state_bytes_per_ion():
    """SoA state read + written once per launch (doubles and ints)."""
    return 2 * (8 * 16 + 4 * 5 + 8 * 4)


def run_gpu(cfg, precision, warmup, steps, window, engine, compact=3):
    from ion_gym_amd.engine import Flight
    model = cfg["model"]
    fp = engine.upload_rz(model.pack(), precision)
    n = len(cfg["ions"]["x"])
    ens = engine.ensemble(n)
    ens.load(cfg["ions"])
    fl = Flight(cfg["dt"], cfg["t_max"], gas=cfg["gas"], max_steps=window,
                compact=compact)
    for _ in range(warmup):
        engine.fly(fp, fl, ens)
    engine.sync()
    s0 = ens.steps()
    kernel_ms = []
    engine.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        engine.fly(fp, fl, ens)
        kernel_ms.append(engine.last_fly_ms())
    engine.sync()
    wall = time.perf_counter() - t0
    s1 = ens.steps()
    live = ens.live()
    n_sets = 1 + len(model.ch_kind)
    lay = fp.layout()
    out = dict(ion_steps=s1 - s0, wall_s=wall, kernel_ms=kernel_ms,
               live_end=live, n=n, n_sets=n_sets,
               stored_sets=lay["n_sets"], field_bytes=fp.nbytes,
               state=ens.read(records=False))
    if precision == "f32":
        out["records"] = stored_records(engine, fp)
    ens.release()
    fp.release()
    return out


def cpu_baseline(cfg, window, n_sample, threads):
    """The oracle (C FP64 port of the reference kernels, OpenMP over ions)
    on a bounded sample of the same workload and window."""
    from oracle import oracle as O
    ions = {k: v[:n_sample] for k, v in cfg["ions"].items()}
    fl = dict(dt=cfg["dt"], t_max=cfg["t_max"], max_steps=window,
              rng=O.RNG_PHILOX, gas=cfg["gas"])
    pack = cfg["model"].pack()
    t0 = time.perf_counter()
    r = O.fly_rz(pack, fl, ions, n_threads=threads)
    el = time.perf_counter() - t0
    return int(r["steps"].sum()), el


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _probe_lib():
    path = os.path.join(HERE, "ion_gym_amd", "libig_probe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    L.igp_gather_rz.argtypes = [ctypes.c_void_p, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_int32,
                                ctypes.POINTER(ctypes.c_double)]
    L.igp_copy.argtypes = [ctypes.c_int64, ctypes.c_int32,
                           ctypes.POINTER(ctypes.c_double)]
    return L


def stored_records(engine, fp):
    """The FP32 r-z record array the flight kernel reads, as stored on
    the device (node-major records of the stored sets' (Ez, Eu) pairs,
    padded to whole 16-byte units; linked drive channels folded)."""
    lay = fp.layout()
    nx, nu = fp.shape
    rec = np.zeros((nx, nu, lay["stride"]), np.float32)
    for s in range(lay["n_sets"]):
        for c in range(2):
            rec[:, :, 2 * s + c] = engine.read_set(fp, s, c)
    return rec


def probe_gather(rec, pack, state, dt, n_iter, reps=3):
    """The gather rate of the flight's own access pattern: each ion of
    ``state`` (slot order, as the r-z flight lays out its lanes) walks
    n_iter sample points from its position with its velocity times dt,
    reading its bilinear cell's four records of ``rec`` (the stored
    records).  Returns (gathers per second, bytes loaded per gather, ms
    per launch)."""
    nx, nu, stride = rec.shape
    h = float(pack["mm"])
    u0 = float(pack["u0"])
    x, y, z = state["x"], state["y"], state["z"]
    r = np.hypot(y, z)
    vr = np.where(r > 0, (y * state["vy"] + z * state["vz"])
                  / np.maximum(r, 1e-300), 0.0)
    pts = np.stack([x / h, (r - u0) / h, state["vx"] * dt / h,
                    vr * dt / h]).astype(np.float32)
    pts = np.ascontiguousarray(pts)
    n = pts.shape[1]
    ms = ctypes.c_double()
    rc = _probe_lib().igp_gather_rz(rec.ctypes.data, nx, nu, stride,
                                    pts.ctypes.data, n, n_iter, reps,
                                    ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"igp_gather_rz failed ({rc})")
    return n * n_iter / (ms.value * 1e-3), 4 * stride * 4, ms.value


COPY_FLOATS = 1 << 28          # 1 GiB of floats: the counters' calibration


def counter_probe_main(args):
    """This script's probe mode, the program a rocprofv3 counter run
    profiles: the throughput workload's first two FP32 windows (the second
    is the measured dispatch), then two plain copies of 1 GiB."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Engine, Flight
    cfg = configs.funnel_rz(n_ions=args.ions, seed=3)
    eng = Engine(0)
    fp = eng.upload_rz(cfg["model"].pack(), "f32")
    ens = eng.ensemble(len(cfg["ions"]["x"]))
    ens.load(cfg["ions"])
    fl = Flight(cfg["dt"], cfg["t_max"], gas=cfg["gas"],
                max_steps=args.window, compact=int(args.compact))
    for _ in range(2):
        eng.fly(fp, fl, ens)
    eng.sync()
    ens.release()
    fp.release()
    eng.close()
    ms = ctypes.c_double()
    rc = _probe_lib().igp_copy(COPY_FLOATS, 2, ctypes.byref(ms))
    print(f"[probe] two windows and two 1 GiB copies done (copy rc {rc}, "
          f"{ms.value:.3f} ms)", flush=True)


def _counter_rows(outdir):
    rows = []
    for f in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"),
                       recursive=True):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _per_dispatch(rows, match):
    """{dispatch id: {counter: value, '_ns': duration}} of the dispatches
    whose kernel name contains ``match``."""
    out = {}
    for r in rows:
        name = r.get("Kernel_Name", "")
        if match not in name:
            continue
        d = out.setdefault(int(r.get("Dispatch_Id", 0)), {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + \
            float(r["Counter_Value"])
        try:
            d["_ns"] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        except (KeyError, TypeError, ValueError):
            pass
    return out


def run_counters(args):
    """Two rocprofv3 counter runs of the probe mode (FETCH_SIZE; then
    WRITE_SIZE with the vector-issue counters), each its own process under
    a time limit.  Returns the flight dispatch's counters per launch and
    the copy calibration, or {'error': ...}."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return {"error": "rocprofv3 not found"}
    passes = (["FETCH_SIZE"],
              ["WRITE_SIZE", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU",
               "SQ_WAVES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"])
    env = dict(os.environ, TMPDIR="/tmp")
    flight, copy = {}, {}
    cmds = []
    for k, counters in enumerate(passes):
        out = tempfile.mkdtemp(prefix="ig_counters_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc", *counters,
               "--output-format", "csv", "-d", out, "-o", "run", "--",
               sys.executable, os.path.join(HERE, "bench.py"),
               "--counter-probe", "--ions", str(args.ions), "--window",
               str(args.window), "--compact", str(args.compact)]
        cmds.append(" ".join(cmd[4:]))
        print(f"[bench] counters pass {k + 1}: {' '.join(counters)}",
              file=sys.stderr, flush=True)
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True,
                           text=True)
        rows = _counter_rows(out)
        shutil.rmtree(out, ignore_errors=True)
        fl = _per_dispatch(rows, "rz_fly_kernel")
        cp = _per_dispatch(rows, "copy_probe_kernel")
        if not fl or not cp:
            return {"error": f"pass {k + 1} ({' '.join(counters)}): rc "
                             f"{r.returncode}, no counter rows for the "
                             f"flight or the copy; "
                             f"{(r.stderr or '')[-300:]}"}
        flight.update(fl[max(fl)])           # the second (timed) window
        for d in cp.values():
            for c, v in d.items():
                copy.setdefault(c, []).append(v)
    copy = {c: float(np.mean(v)) for c, v in copy.items()}
    return {"flight": flight, "copy": copy, "commands": cmds}


def funnel_precision(eng, cfg, steps=4000):
    """FP32 basis against FP64 on the throughput workload itself: the same
    births, one window of ``steps`` steps in each precision; two-sample KS
    p-values and mean ratios of the axial displacement, final radius,
    kinetic energy and collision count, and the fate fractions."""
    from scipy import stats
    from ion_gym_amd.engine import Flight
    res = {}
    for prec in ("f32", "f64"):
        fp = eng.upload_rz(cfg["model"].pack(), prec)
        n = len(cfg["ions"]["x"])
        ens = eng.ensemble(n)
        ens.load(cfg["ions"])
        eng.fly(fp, Flight(cfg["dt"], cfg["t_max"], gas=cfg["gas"],
                           max_steps=steps), ens)
        g = ens.read(records=False)
        ens.release()
        fp.release()
        fly = g["fate"] == -2
        m = cfg["ions"]["m_amu"] * 1.66053906660e-27
        ke = 0.5 * m * (g["vx"] ** 2 + g["vy"] ** 2 + g["vz"] ** 2) * 1e6 \
            / 1.602176634e-19
        res[prec] = dict(dx=(g["x"] - cfg["ions"]["x"])[fly],
                         rad=np.hypot(g["y"], g["z"])[fly], ke=ke[fly],
                         ncol=g["ncol"][fly].astype(float),
                         flying=float(fly.mean()),
                         splat=float((g["fate"] == 0).mean()))
    a, b = res["f32"], res["f64"]
    out = {"workload": "funnel_rz, the throughput workload's births",
           "ions": int(len(cfg["ions"]["x"])), "window_steps": steps,
           "flying_f32": a["flying"], "flying_f64": b["flying"],
           "splat_f32": a["splat"], "splat_f64": b["splat"]}
    for k in ("dx", "rad", "ke", "ncol"):
        out[f"ks_p_{k}"] = float(stats.ks_2samp(a[k], b[k]).pvalue)
        out[f"mean_ratio_{k}"] = float(a[k].mean() / b[k].mean())
    return out

def run_serving(eng, n_req, loads, flight_us, window, capacity, seed,
                compact=3):
    """The serving leg: requests on the SLIM traveling-wave board (11 FP32
    field sets, N2 4 Torr hard sphere) arrive as a Poisson stream and are
    served by one FlightServer (continuous batching over one resident
    slab).  For each offered load the arrival rate is that fraction of the
    engine's measured solo throughput divided by the mean request work;
    the same n_req requests are replayed at every load.  Returns, per
    load, p50/p99 request latency (submit -> result on the host) and the
    sustained ion-steps/s over the run."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    from ion_gym_amd.serve import (FlightRequest, FlightServer, latency_stats,
                                   make_tune)
    import ion_gym_amd._lib as L
    from ion_gym_amd.serve import _fill_tune
    t0 = time.perf_counter()
    c = configs.slim_3d(coarse=1)
    fp = eng.upload_xyz(c["pack"], "f32")
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    specs, work = [], []
    for k in range(n_req):
        ions, tune = configs.slim_request(c, rng, t_flight_us=flight_us,
                                          rng_key=k + 1)
        t = make_tune(c["pack"], **tune)
        specs.append((ions, t))
        work.append(len(ions["x"]) * math.ceil(t["t_max"] / t["dt"]))
    # solo throughput: one 200,000-ion request, two windows
    cal_ions, cal_tune = configs.slim_request(c, np.random.default_rng(seed + 1),
                                              n_ions=200_000,
                                              t_flight_us=flight_us)
    ens = eng.ensemble(200_000)
    tunes = (L.Tune * 1)()
    _fill_tune(tunes[0], make_tune(c["pack"], **cal_tune))
    L.check(L.lib().ig_ensemble_set_tunes(eng.h, ens.id, 1, tunes))
    ens.load(cal_ions)
    fl = Flight(cal_tune["dt"], cal_tune["t_max"], gas=c["gas"],
                max_steps=window, compact=compact)
    eng.fly(fp, fl, ens)
    s0 = ens.steps()
    t1 = time.perf_counter()
    eng.fly(fp, fl, ens)
    eng.sync()
    solo = (ens.steps() - s0) / (time.perf_counter() - t1)
    ens.release()
    mean_work = float(np.mean(work))
    per_load = []
    for li, load in enumerate(loads):
        lam = load * solo / mean_work
        gaps = np.random.default_rng(seed + 2 + li).exponential(1.0 / lam,
                                                                n_req)
        arrivals = [(float(t), FlightRequest(ions=ions, tune=tune))
                    for t, (ions, tune) in zip(np.cumsum(gaps) - gaps[0],
                                               specs)]
        srv = FlightServer(eng, fp, Flight(cal_tune["dt"], cal_tune["t_max"],
                                           gas=c["gas"], compact=compact),
                           capacity=capacity, window=window)

        def progress(n_done, load=load):
            print(f"[bench] serving at load {load}: {n_done}/{n_req} "
                  f"requests done", file=sys.stderr, flush=True)

        t2 = time.perf_counter()
        done = srv.run_arrivals(arrivals, progress=progress)
        wall = time.perf_counter() - t2
        st = latency_stats(done)
        per_load.append({"offered_load": load, "arrival_rate_per_s": lam,
                         "requests": len(done), "p50_s": st["p50_s"],
                         "p99_s": st["p99_s"], "mean_s": st["mean_s"],
                         "sustained_ion_steps_per_s": srv.ion_steps / wall,
                         "wall_s": wall})
        srv.release()
    out = {"workload": "slim_3d (SLIM TW board 201x81x801 @ 0.05 mm, 11 "
                       "FP32 sets, N2 4 Torr HS)",
           "requests_per_load": n_req, "ions_per_request": "log-U(1e3, 5e5)",
           "flight_us": flight_us, "window_steps": window,
           "slab_capacity": capacity, "compact": compact,
           "solo_ion_steps_per_s": solo, "loads": per_load,
           "field_bytes": fp.nbytes, "build_s": build_s}
    fp.release()
    return out


def run_precision(eng, n_ions, n_cycles):
    """The FP32-basis error on the 3-D quadrupole at its full grid
    (161 x 161 x 401 @ 0.05 mm, RF 1 MHz 200 V0-p with DC at a = 0.02,
    m/z 400-600, vacuum): the same births flown with FP64 and FP32-basis
    packs; transmission (exit through the far end plane) of each and the
    per-ion position and time-of-flight differences of ions that share
    their fate."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.quad_3d(n_ions=n_ions, n_cycles=n_cycles)
    z_end = (c["pack"]["EAx"].shape[2] - 1) * c["pack"]["h_mm"]
    out = {}
    res = {}
    for prec in ("f64", "f32"):
        fp = eng.upload_xyz(c["pack"], prec)
        t0 = time.perf_counter()
        g = eng.fly_batch(fp, Flight(c["dt"], c["t_max"]), c["ions"])
        el = time.perf_counter() - t0
        fp.release()
        res[prec] = g
        out[f"{prec}_ion_steps_per_s"] = float(g["steps"].sum()) / el
        out[f"transmission_{prec}"] = float(np.mean(
            (g["fate"] == 1) & (g["z"] >= z_end - 1e-9)))
    a, b = res["f32"], res["f64"]
    same = a["fate"] == b["fate"]
    d = np.sqrt((a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 +
                (a["z"] - b["z"]) ** 2)[same]
    dt_rel = (np.abs(a["t_us"] - b["t_us"]) / np.abs(b["t_us"]))[same]
    out.update({
        "workload": "quad_3d (161x161x401 @ 0.05 mm, 3 sets, vacuum)",
        "ions": int(n_ions), "rf_cycles": int(n_cycles),
        "same_fate_fraction": float(same.mean()),
        "pos_err_mm_median": float(np.median(d)),
        "pos_err_mm_p99": float(np.percentile(d, 99)),
        "pos_err_mm_max": float(d.max()),
        "tof_rel_err_median": float(np.median(dt_rel)),
        "tof_rel_err_max": float(dt_rel.max())})
    return out


def run_configs(eng):
    """The other flight configurations, each measured as ion-steps/s in
    FP64 and in the FP32 basis: the einzel lens (10,000 ions in vacuum,
    flown to fate in one launch), the batched drift tube (64 SDS ensembles
    of log-uniform 1e3-2e5 ions, each its own field strength, m/z and dt,
    in one slab: one warm-up window, then two timed 500-step windows), the
    planar route's quadrupole guide and the 3-D SDS drift cell (1M ions
    each, a warm-up window and a timed one)."""
    from ion_gym_amd import configs
    from ion_gym_amd.batch import EnsembleBatch
    from ion_gym_amd.engine import Flight
    out = {}
    c = configs.einzel_rz(n_ions=10_000)
    res = {}
    for prec in ("f64", "f32"):
        fp = eng.upload_rz(c["model"].pack(), prec)
        # a short flight of a few ions first: the kernel's code is loaded
        # on its first launch
        warm = eng.ensemble(64)
        warm.load({k: v[:64] for k, v in c["ions"].items()})
        eng.fly(fp, Flight(c["dt"], c["t_max"], max_steps=10), warm)
        eng.sync()
        warm.release()
        ens = eng.ensemble(len(c["ions"]["x"]))
        ens.load(c["ions"])
        eng.sync()
        t0 = time.perf_counter()
        eng.fly(fp, Flight(c["dt"], c["t_max"], max_steps=0), ens)
        eng.sync()
        el = time.perf_counter() - t0
        res[prec] = ens.steps() / el
        out[f"einzel_live_end_{prec}"] = ens.live()
        ens.release()
        fp.release()
    out["einzel"] = {"workload": "einzel_rz (401x101 @ 0.1 mm, 0/+600/0 V, "
                                 "10,000 ions m/z 100, vacuum, to fate)",
                     "f64_ion_steps_per_s": res["f64"],
                     "f32_ion_steps_per_s": res["f32"]}
    b = configs.drift_batch_rz()
    res = {}
    for prec in ("f64", "f32"):
        fp = eng.upload_rz(b["model"].pack(), prec)
        fl = Flight(1.0, 2000.0, gas=b["gas"], max_steps=500)
        batch = EnsembleBatch(eng, fp, fl, [(e["ions"], e["tune"])
                                            for e in b["ensembles"]])
        batch.fly(1, graph=False)
        eng.sync()
        s0 = batch.slab.steps()
        t0 = time.perf_counter()
        batch.fly(2, graph=False)
        eng.sync()
        el = time.perf_counter() - t0
        res[prec] = (batch.slab.steps() - s0) / el
        n = batch.n
        batch.release()
        fp.release()
    out["drift_batch"] = {"workload": "drift_batch_rz (801x101 @ 0.1 mm, N2 "
                                      "4 Torr SDS, 64 ensembles)",
                          "ions": n, "window_steps": 500,
                          "f64_ion_steps_per_s": res["f64"],
                          "f32_ion_steps_per_s": res["f32"]}
    # the planar route and the 3-D SDS route: 1M ions, one warm-up window,
    # one timed window
    pq = configs.planar_quad(n_ions=1_000_000)
    d3 = configs.drift_3d(n_ions=1_000_000)
    legs = (("planar_quad", eng.upload_planar, pq["pack"], pq["ions"],
             Flight(pq["dt"], pq["t_max"], max_steps=2000),
             "planar_quad (161x161 @ 0.05 mm linear quadrupole guide, RF "
             "1 MHz, vacuum, 1M ions)", 2000),
            ("drift_3d", eng.upload_xyz, d3["pack"], d3["ions"],
             Flight(d3["dt"], d3["t_max"], gas=d3["gas"], planes=d3["planes"],
                    max_steps=1000),
             "drift_3d (41x41x301 @ 0.1 mm SDS drift cell, N2 4 Torr, 1M "
             "ions)", 1000))
    for name, upload, pack, ions, fl, desc, win in legs:
        res = {}
        for prec in ("f64", "f32"):
            fp = upload(pack, prec)
            ens = eng.ensemble(len(ions["x"]))
            ens.load(ions)
            eng.fly(fp, fl, ens)
            eng.sync()
            s0 = ens.steps()
            t0 = time.perf_counter()
            eng.fly(fp, fl, ens)
            eng.sync()
            res[prec] = (ens.steps() - s0) / (time.perf_counter() - t0)
            ens.release()
            fp.release()
        out[name] = {"workload": desc, "window_steps": win,
                     "f64_ion_steps_per_s": res["f64"],
                     "f32_ion_steps_per_s": res["f32"]}
    return out


def solver_scene(shape):
    """One electrode's basis: a 3-node-thick plate at 1e4 V across the
    middle half of x and three quarters of z, halfway up y, between
    grounded walls at y = 0 and y = ny - 1 (the scene of the reference's
    multigrid gates, tests/test_multigrid.py, at board size)."""
    nx, ny, nz = shape
    fixed = np.zeros(shape, bool)
    fixed[nx // 4:3 * nx // 4, ny // 2 - 1:ny // 2 + 2,
          nz // 8:7 * nz // 8] = True
    val = np.zeros(shape)
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    fixed[:, -1, :] = True
    return fixed, val


def run_solver(eng, shape, cpu_shape, cpu):
    """The basis solve leg: one V-cycle solve and one cascade solve to 1e-6
    V of the scene above on a SLIM-sized grid, the red-black SOR pair time (the memory-bound
    stencil) from the difference of a 1000- and a 200-pair run, and, with
    cpu, the oracle's V-cycle on a smaller grid beside the engine's on the
    same grid (bitwise compared)."""
    from ion_gym_amd.physics import solver3d as S
    fixed, val = solver_scene(shape)
    n = fixed.size
    S.solve3d_vmg(fixed, val, tol=1e-6, engine=eng)
    t0 = time.perf_counter()
    phi, sw, d = S.solve3d_vmg(fixed, val, tol=1e-6, engine=eng)
    gpu_s = time.perf_counter() - t0
    # the cascade (solve3d_mg) on the same basis, warm
    S.solve3d_mg(fixed, val, tol=1e-6, engine=eng)
    t0 = time.perf_counter()
    _, sw_c, d_c = S.solve3d_mg(fixed, val, tol=1e-6, engine=eng)
    cas_s = time.perf_counter() - t0
    runs = {}
    for pairs in (200, 1000):
        t0 = time.perf_counter()
        try:
            S.solve3d(fixed, val, tol=1e-12, max_sweeps=pairs, engine=eng)
        except S.SolveNotConverged:
            pass
        runs[pairs] = time.perf_counter() - t0
    pair_s = (runs[1000] - runs[200]) / 800
    # per red-black pair: both colours read every potential and the mask
    # and write their half of the potentials
    pair_bytes = 2 * n * (8 + 1 + 4)
    out = {"workload": f"V-cycle basis solve {shape[0]}x{shape[1]}x"
                       f"{shape[2]} (one plate at 1e4 V, grounded walls)",
           "nodes": n, "tol_v": 1e-6, "fine_sweep_pairs": sw,
           "cycle_delta_v": d, "gpu_s": gpu_s,
           "cascade": {"gpu_s": cas_s, "fine_sweep_pairs": sw_c,
                       "delta_v": d_c},
           "sor_pair_us": pair_s * 1e6,
           "sor_roofline": {"bound": "hbm", "achieved": pair_bytes / pair_s
                            / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": pair_bytes / pair_s / 1e9
                            / HBM_PEAK_GBS, "bytes_per_pair": pair_bytes}}
    if cpu:
        from oracle import solve3d as O
        fc, vc = solver_scene(cpu_shape)
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
        t0 = time.perf_counter()
        pc, swc, dc, st = O.solve3d_vmg(fc, vc, tol=1e-6,
                                        sym_axes=S.symmetric_axes(fc, vc))
        cpu_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        pg, swg, dg = S.solve3d_vmg(fc, vc, tol=1e-6, engine=eng)
        gpu_small = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": cpu_s, "unit": "s per basis", "cores": threads,
            "kind": "port",
            "sample": f"the same scene at {cpu_shape[0]}x{cpu_shape[1]}x"
                      f"{cpu_shape[2]}, V-cycle to 1e-6 V (FP64, OpenMP)"}
        out["gpu_s_on_cpu_sample"] = gpu_small
        out["speedup_vs_cpu"] = cpu_s / gpu_small
        out["bitwise_equal_to_cpu"] = bool(st == 0 and swc == swg
                                           and np.array_equal(pc, pg))
    out["rz_2d"] = run_solver_2d(eng, cpu)
    return out


def run_solver_2d(eng, cpu):
    """The 2-D leg: the funnel's 1201 x 201 r-z grid (100 rings on an
    axial ladder of 100 -> 1 V, open ends and outer radius), device SOR
    to 1e-12 of the scale, warm; with cpu, the oracle's sparse LU of the
    same system on the host and the largest difference."""
    from ion_gym_amd.configs import _grid_labels
    from ion_gym_amd.physics import solver2d as S2
    h = 0.1
    elec = [(i + 1, 10.0 + i, 10.5 + i, 12.5 - 11.5 * i / 99.0,
             15.5 - 11.5 * i / 99.0) for i in range(100)]
    lab = _grid_labels(1201, 201, h, elec)
    val = np.where(lab > 0, 100.0 - (lab - 1) * 1.0, 0.0)
    edges = ("z0", "z1", "u1")
    solve = S2.laplace_factor(lab > 0, h, "cylindrical", edges, engine=eng)
    solve(val)
    t0 = time.perf_counter()
    p = solve(val)
    out = {"workload": "r-z basis solve 1201x201 (funnel rings, ladder "
                       "100 -> 1 V)", "gpu_s": time.perf_counter() - t0,
           "sor_pairs": solve.sweeps}
    if cpu:
        from oracle import solve2d as O2
        t0 = time.perf_counter()
        q, _ = O2.solve(lab > 0, val, h, "cylindrical", edges)
        out["cpu_lu_s"] = time.perf_counter() - t0
        out["max_abs_diff_v"] = float(np.abs(p - q).max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--window", type=int, default=2000,
                    help="integrator steps per ion per benchmark step")
    ap.add_argument("--ions", type=int, default=1_000_000)
    ap.add_argument("--precision", default="f32",
                    choices=["f32", "f32_store", "f64"])
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the plain FP64 comparison run")
    ap.add_argument("--cpu-sample", type=int, default=32000)
    ap.add_argument("--cpu-window", type=int, default=40000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compact", type=int, default=3, choices=(0, 1, 2, 3),
                    help="live-ion lane list before each window: 0 none, 1 "
                         "slot order, 2 cell order, 3 the engine's choice "
                         "per route (results are identical either way)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32-vs-FP64 quadrupole error leg")
    ap.add_argument("--precision-ions", type=int, default=100_000)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving leg (SLIM board, Poisson arrivals)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests per offered load")
    ap.add_argument("--serve-loads", default="0.3,0.6,0.9",
                    help="offered loads (fractions of the solo throughput)")
    ap.add_argument("--serve-flight-us", type=float, default=2.0)
    ap.add_argument("--serve-window", type=int, default=500)
    ap.add_argument("--serve-capacity", type=int, default=2_000_000)
    ap.add_argument("--no-solver", action="store_true",
                    help="skip the 3-D Laplace basis-solve leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the einzel, batched drift-tube, planar and "
                         "3-D SDS legs")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter runs (HBM traffic, "
                         "vector issue)")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the gather-rate probe")
    ap.add_argument("--counter-probe", action="store_true",
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.counter_probe:
        counter_probe_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not "
              f"supported (across GPUs the ions would shard, one process per "
              f"GPU, each with its own copy of the field pack)",
              file=sys.stderr)
        sys.exit(2)

    from ion_gym_amd import configs
    from ion_gym_amd.engine import Engine
    print(f"[bench] building funnel_rz with {args.ions} ions", file=sys.stderr,
          flush=True)
    cfg = configs.funnel_rz(n_ions=args.ions, seed=3)
    eng = Engine(0)
    print(f"[bench] timing {args.precision}: warmup {args.warmup}, steps "
          f"{args.steps}, window {args.window}", file=sys.stderr, flush=True)
    main_run = run_gpu(cfg, args.precision, args.warmup, args.steps,
                       args.window, eng, int(args.compact))
    value = main_run["ion_steps"] / main_run["wall_s"]
    ms_per_step = main_run["wall_s"] / args.steps * 1e3
    fp64 = None
    if not args.no_fp64 and args.precision != "f64":
        print("[bench] timing the plain FP64 kernel", file=sys.stderr,
              flush=True)
        r64 = run_gpu(cfg, "f64", 1, max(2, args.steps // 4), args.window,
                      eng, int(args.compact))
        fp64 = r64["ion_steps"] / r64["wall_s"]
    # roofline of the flight kernel
    vb = 8 if args.precision == "f64" else 4
    per_step = gather_bytes_per_ion_step(main_run["n_sets"], vb)
    launches = args.steps
    ion_steps_per_launch = main_run["ion_steps"] / launches
    alg_bytes = ion_steps_per_launch * per_step + main_run["n"] * \
        state_bytes_per_ion()
    mean_ms = float(np.mean(main_run["kernel_ms"]))
    kernel_rate = ion_steps_per_launch / (mean_ms * 1e-3)  # gathers/s
    roof = {"kernel": "rz_fly_kernel", "kernel_ms": mean_ms,
            "alg_bytes_per_launch": alg_bytes,
            "alg_bytes_per_ion_step": per_step,
            "alg_gbs": alg_bytes / (mean_ms * 1e-3) / 1e9,
            # what the kernel itself gathers per ion-step: one bilinear
            # evaluation of the stored sets (the step's second evaluation
            # is the next step's first) and the metal byte
            "stored_sets": main_run["stored_sets"],
            "kernel_bytes_per_ion_step":
                4 * 2 * main_run["stored_sets"] * vb + 1}
    if not args.no_probe and args.precision == "f32":
        print("[bench] gather-rate probe: the ions' own sample points "
              "through the FP32 records", file=sys.stderr, flush=True)
        p_rate, p_bytes, p_ms = probe_gather(main_run["records"],
                                             cfg["model"].pack(),
                                             main_run["state"], cfg["dt"],
                                             args.window)
        gather = dict(bound="l2_mall", unit="GB/s",
                      achieved=kernel_rate * p_bytes / 1e9,
                      peak=p_rate * p_bytes / 1e9,
                      frac=kernel_rate / p_rate,
                      bytes_per_gather=p_bytes, probe_ms=p_ms,
                      peak_source="probe_gather.hip: every ion's own "
                                  "sample-point walk through the stored "
                                  "records, one gather per step, nothing "
                                  "else")
        roof["gather"] = gather
        # without the counter runs the gather floor is the stated bound
        roof.update({k: gather[k] for k in ("bound", "unit", "achieved",
                                            "peak", "frac")})
    counters = None
    if not args.no_counters:
        counters = run_counters(args)
    traffic = None
    if counters and "flight" in counters:
        fl, cp = counters["flight"], counters["copy"]
        kib = COPY_FLOATS * 4 / 1024.0
        f_fac = kib / cp["FETCH_SIZE"]
        w_fac = kib / cp["WRITE_SIZE"]
        traffic = int(round((fl["FETCH_SIZE"] * f_fac
                             + fl["WRITE_SIZE"] * w_fac) * 1024))
        busy_cycles = fl["SQ_ACTIVE_INST_VALU"] * 4.0 / SIMDS
        kern_s = fl.get("_ns", mean_ms * 1e6) * 1e-9
        waves = fl["SQ_WAVES"]
        roof.update(
            traffic=traffic,
            hbm={"achieved": traffic / (mean_ms * 1e-3) / 1e9,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": traffic / (mean_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            valu_busy=busy_cycles / (CLOCK_HZ * kern_s),
            valu_insts_per_wave_step=fl["SQ_INSTS_VALU"] / waves
            / args.window,
            counter_calibration={"fetch_factor": f_fac,
                                 "write_factor": w_fac,
                                 "copy_kib": kib},
            counter_kernel_ms=kern_s * 1e3)
        # vector issue: a wave64 VALU instruction holds a 16-lane SIMD for
        # four cycles, so the chip issues at most SIMDS * clock / 4
        valu_rate = fl["SQ_INSTS_VALU"] / kern_s
        valu_peak = SIMDS * CLOCK_HZ / 4.0
        roof["valu"] = {"achieved": valu_rate / 1e9, "peak": valu_peak / 1e9,
                        "unit": "G wave-instructions/s",
                        "frac": valu_rate / valu_peak,
                        "busy": roof["valu_busy"]}
        g_frac = roof.get("gather", {}).get("frac", 0.0)
        roof["binding"] = ("valu_issue" if roof["valu"]["frac"] > g_frac
                           else "l2_mall")
        if roof["binding"] == "valu_issue":
            roof.update(bound="valu_issue", unit="G wave-instructions/s",
                        achieved=valu_rate / 1e9, peak=valu_peak / 1e9,
                        frac=valu_rate / valu_peak)
    elif counters:
        roof["traffic"] = None
        roof["counters_error"] = counters["error"]
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
        print(f"[bench] CPU baseline: {args.cpu_sample} ions x "
              f"{args.cpu_window} steps on {threads} threads",
              file=sys.stderr, flush=True)
        st, el = cpu_baseline(cfg, args.cpu_window, args.cpu_sample, threads)
        ns = threads / (st / el) * 1e9
        cpu = {"value": st / el, "unit": "ion-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_sample} ions of the same workload x "
                         f"{args.cpu_window} steps (FP64, HS, Philox), "
                         f"{el:.1f} s",
               "cpu_model": cpu_model(),
               "ns_per_ion_step_per_core": ns,
               "reference_ns_per_ion_step_per_core": list(REF_NS_PER_ION_STEP),
               "ratio_to_reference": [ns / r for r in REF_NS_PER_ION_STEP],
               "gpu_over_cpu": value / (st / el),
               "gpu_over_reference_cores": [
                   value / (threads / (r * 1e-9))
                   for r in REF_NS_PER_ION_STEP]}
    precision_funnel = None
    if not args.no_precision:
        print("[bench] precision on the funnel: FP32 basis vs FP64, "
              "4000-step window", file=sys.stderr, flush=True)
        precision_funnel = funnel_precision(eng, cfg)
    precision = None
    if not args.no_precision:
        print(f"[bench] precision: quad_3d full grid, "
              f"{args.precision_ions} ions, FP32 basis vs FP64",
              file=sys.stderr, flush=True)
        precision = run_precision(eng, args.precision_ions, 2000)
    serving = None
    if not args.no_serve:
        loads = [float(v) for v in args.serve_loads.split(",") if v]
        print(f"[bench] serving: {args.serve_requests} SLIM requests at "
              f"each offered load {loads}", file=sys.stderr, flush=True)
        serving = run_serving(eng, args.serve_requests, loads,
                              args.serve_flight_us, args.serve_window,
                              args.serve_capacity, seed=11,
                              compact=int(args.compact))
    others = None
    if not args.no_configs:
        print("[bench] configurations: einzel lens, batched drift tube, "
              "planar quadrupole, 3-D SDS drift cell",
              file=sys.stderr, flush=True)
        others = run_configs(eng)
    solver = None
    if not args.no_solver:
        print("[bench] solver: V-cycle basis solve 201x81x801",
              file=sys.stderr, flush=True)
        solver = run_solver(eng, (201, 81, 801), (101, 41, 401),
                            not args.no_cpu)
    line = {
        "metric": "ion-steps/s at 1M ions (3-D funnel+HS), FP32 vs FP64 "
                  "plain; p50/p99 latency",
        "value": value,
        "unit": "ion-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision != "f64" else "f64",
        "data": "synthetic",
        "config": {"workload": "funnel_rz (ion funnel r-z + hard-sphere He)",
                   "ions": main_run["n"], "grid": "1201x201 @ 0.1 mm "
                   "(unfolded 1201x401)", "window_steps": args.window,
                   "field_precision": args.precision,
                   "state_precision": "f64", "dt_ns": cfg["dt"] * 1e3,
                   "live_ions_end": main_run["live_end"],
                   "compact": int(args.compact),
                   "field_bytes": main_run["field_bytes"]},
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": (value / fp64) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "precision_funnel": precision_funnel,
        "serving": serving,
        "precision": precision,
        "solver": solver,
        "configs": others,
        "design_notes": DESIGN_NOTES,
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
