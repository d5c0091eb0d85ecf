"""Launch settings change no result: the same funnel ions (hard-sphere He,
FP64 and FP32 basis) and SLIM ions flown under different block sizes,
register-budget variants and collision batches end bit for bit alike; the
autotuner measures real launches, picks one of its candidates and caches
it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "ncol", "steps")
SETTINGS = [dict(block=256, occupancy=1, coll_batch=1),
            dict(block=64, occupancy=2, coll_batch=16),
            dict(block=192, occupancy=3, coll_batch=32),
            dict(block=128, occupancy=3, coll_batch=4)]


def _fly(engine, fp, fl, ions, steps):
    ens = engine.ensemble(len(ions["x"]))
    ens.load(ions)
    f = fl.launch(**{})
    f.max_steps = steps
    engine.fly(fp, f, ens)
    g = ens.read()
    ens.release()
    return g


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_funnel_results_do_not_depend_on_launch_settings(engine, prec):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.funnel_rz(n_ions=3000, seed=3, coarse=2)
    fp = engine.upload_rz(c["model"].pack(), prec)
    base = Flight(c["dt"], c["t_max"], gas=c["gas"])
    outs = [_fly(engine, fp, base.launch(**s), c["ions"], 1500)
            for s in SETTINGS]
    assert outs[0]["ncol"].sum() > 0
    for o in outs[1:]:
        for k in KEYS:
            assert np.array_equal(o[k], outs[0][k]), k
    fp.release()


def test_slim_results_do_not_depend_on_launch_settings(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.slim_3d(coarse=4)
    rng = np.random.default_rng(4)
    ions, tune = configs.slim_request(c, rng, n_ions=2000, t_flight_us=0.2)
    pk = dict(c["pack"])
    for k in ("ch_om", "ch_amp", "ch_off"):
        pk[k] = np.asarray(tune[k], float)
    fp = engine.upload_xyz(pk, "f32")
    base = Flight(tune["dt"], tune["t_max"], gas=c["gas"])
    outs = [_fly(engine, fp, base.launch(**s), ions, 600) for s in SETTINGS]
    assert outs[0]["ncol"].sum() > 0
    for o in outs[1:]:
        for k in KEYS:
            assert np.array_equal(o[k], outs[0][k]), k
    fp.release()


def test_autotuner_measures_and_caches(engine, tmp_path):
    from ion_gym_amd import autotune as AT
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.funnel_rz(n_ions=20000, seed=3, coarse=2)
    fp = engine.upload_rz(c["model"].pack(), "f32")
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"])
    path = str(tmp_path / "launch.json")
    r = AT.autotune(engine, fp, fl, c["ions"], sample=20000, window=100,
                    reps=1, cache_path=path)
    cands = AT.candidates("f32", fl.gas["model"], fp.route)
    assert {k: r[k] for k in ("block", "occupancy", "coll_batch")} in cands
    assert len(r["table"]) == len(cands)
    assert all(row["ms"] > 0 for row in r["table"])
    assert r["ms"] == min(row["ms"] for row in r["table"])
    again = AT.autotune(engine, fp, fl, c["ions"], cache_path=path)
    assert again["cached"]
    g = _fly(engine, fp, AT.tuned(fl, again), c["ions"], 300)
    assert np.all(g["steps"] == 300)
    fp.release()


def test_vacuum_sds_and_planar_results_do_not_depend_on_launch_settings(
        engine):
    """Every register-budget variant of the vacuum and SDS kernels (r-z
    einzel lens, r-z SDS drift tube, FP64 and FP32 basis) and the planar
    guide's FP32 vacuum kernel at 1-4 waves per SIMD."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    cases = []
    e = configs.einzel_rz(n_ions=3000, coarse=2)
    cases.append(("einzel", e["model"].pack(), e, None, engine.upload_rz,
                  (1, 2, 3)))
    d = configs.drift_rz(n_ions=3000, coarse=2)
    dpk = dict(d["model"].pack())
    dpk["EzA"] = dpk["EzA"] * d["scale"]
    dpk["EuA"] = dpk["EuA"] * d["scale"]
    cases.append(("drift", dpk, d, d["gas"], engine.upload_rz, (1, 2, 3)))
    q = configs.planar_quad(n_ions=3000, coarse=2)
    cases.append(("planar", q["pack"], q, None, engine.upload_planar,
                  (1, 2, 3, 4)))
    for name, pk, c, gas, upload, occs in cases:
        for prec in ("f64", "f32"):
            print(f"launch invariance {name} {prec}", flush=True)
            fp = upload(pk, prec)
            base = Flight(c["dt"], c["t_max"], gas=gas)
            outs = [_fly(engine, fp, base.launch(occupancy=o, block=b),
                         c["ions"], 800)
                    for o, b in zip(occs, (256, 64, 192, 128))]
            for o in outs[1:]:
                for k in KEYS:
                    assert np.array_equal(o[k], outs[0][k]), (name, prec, k)
            fp.release()
