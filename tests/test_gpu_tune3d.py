"""A tuned 3-D flight (the drive read from the ion's tune: packed per-set
records, FP32-basis weights) equals the untuned flight of the same pack
bit for bit when the tune repeats the pack's drive at set scale 1: for
sin, cos, square at duty 0.5 and square at another duty (the general
path), on the 3-set quadrupole in vacuum and the 11-set SLIM board with
hard-sphere gas."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "ncol", "steps")


def _fly_both(engine, pack, ions, dt, t_max, gas, prec, steps):
    from ion_gym_amd import _lib as L
    from ion_gym_amd.engine import Flight
    from ion_gym_amd.serve import _fill_tune, make_tune
    fp = engine.upload_xyz(pack, prec)
    n = len(ions["x"])
    out = []
    for tuned in (False, True):
        ens = engine.ensemble(n)
        if tuned:
            tunes = (L.Tune * 1)()
            _fill_tune(tunes[0], make_tune(pack, dt=dt, t_max=t_max))
            L.check(L.lib().ig_ensemble_set_tunes(engine.h, ens.id, 1, tunes))
        ens.load(ions)
        engine.fly(fp, Flight(dt, t_max, gas=gas, max_steps=steps), ens)
        out.append(ens.read(records=False))
        ens.release()
    fp.release()
    return out


@pytest.mark.parametrize("kinds", [((0, 0.5), (2, 0.3)), ((1, 0.5), (2, 0.5))])
def test_quad_tuned_equals_untuned(engine, kinds):
    from ion_gym_amd import configs
    c = configs.quad_3d(n_ions=2000, coarse=4, n_cycles=40)
    pack = dict(c["pack"])
    pack["ch_kind"] = np.array([k for k, _ in kinds], np.int64)
    pack["ch_duty"] = np.array([d for _, d in kinds], float)
    print(f"quad tuned vs untuned, kinds {kinds}", flush=True)
    for prec in ("f32", "f64"):
        a, b = _fly_both(engine, pack, c["ions"], c["dt"], c["t_max"], None,
                         prec, 3000)
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), (prec, k)
        assert a["steps"].min() > 0


def test_slim_tuned_equals_untuned(engine):
    from ion_gym_amd import configs
    c = configs.slim_3d(coarse=4, workers=4)
    ions, tune = configs.slim_request(c, np.random.default_rng(3),
                                      n_ions=3000)
    pack = dict(c["pack"])
    # the request's own drive as the pack's, so the tune repeats it
    for key in ("ch_om", "ch_ph", "ch_amp", "ch_off", "ch_duty"):
        pack[key] = np.asarray(tune[key], float)
    pack["ch_kind"] = np.asarray(tune["ch_kind"], np.int64)
    print("SLIM coarse 4 tuned vs untuned", flush=True)
    a, b = _fly_both(engine, pack, ions, tune["dt"], tune["t_max"], c["gas"],
                     "f32", 1500)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert a["ncol"].sum() > 0
